"""Oracle pinning of the a-posteriori error indicators
(mpet/adaptivempetsolver.py:82-185) by closed-form known answers:

* a P2-exact quadratic displacement u = (x^2, 0, 0): the cell residual is
  div sigma(u) = (2 (2 mu + lambda), 0, 0) everywhere and every jump
  vanishes, so zeta_u = h^2 |K| (2 (2 mu + lambda))^2;
* a displacement and a pressure with a kink across the shared facet of two
  cells (u = max(0, s) e, p = max(0, s), s the signed distance): the facet
  jumps are sigma(s e) n = mu e + (mu + lambda)(e . n) n and K, so each
  cell gets half of avg(h) |F| |RE|^2, and the only cell residual is
  -alpha_1 grad p_1 on the side where p_1 = s;
* a linear pressure switched on over one step with exchange S: the time
  and exchange terms of RK_p, the grad p term of RK_u and R4p against
  exact tetrahedron moments."""
import numpy as np

from oracle import mpet_oracle as O


def _two_cells():
    from waterscapes_amd.fem import MixedSpace
    from waterscapes_amd.mesh import Mesh
    X = np.array([[0.1, 0.2, 0.9], [0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.2, 1.1, 0.1],
                  [0.5, 0.4, -0.8]])
    mesh = Mesh(X, np.array([[0, 1, 2, 3], [1, 2, 3, 4]]))
    space = MixedSpace(mesh, 2, total_pressure=False)
    from conftest import oracle_mesh
    return space, oracle_mesh(space)


def _box(n=2):
    from waterscapes_amd.fem import MixedSpace
    from waterscapes_amd.mesh import BoxMesh
    from conftest import oracle_mesh
    mesh = BoxMesh((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), n, n, n)
    space = MixedSpace(mesh, 2, total_pressure=False)
    return space, oracle_mesh(space)


PARAMS = dict(J=2, E=800.0, nu=0.3, alpha=(0.7, 0.4), c=(1e-2, 2e-2), K=(0.3, 0.05),
              S=((0.0, 0.0), (0.0, 0.0)))


def _h_vol(m):
    v = m["coords"][m["cells"]]
    h = np.max([np.linalg.norm(v[:, i] - v[:, j], axis=1)
                for i in range(4) for j in range(i + 1, 4)], axis=0)
    vol = np.abs(np.linalg.det(np.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0],
                                         v[:, 3] - v[:, 0]], 2))) / 6
    return h, vol


def test_quadratic_displacement_cell_residual():
    space, m = _box()
    x = np.zeros(space.ndofs)
    X2 = m["x2"]
    x[: space.nu] = np.stack([X2[:, 0] ** 2, 0 * X2[:, 0], 0 * X2[:, 0]], 1).reshape(-1)
    zp, zu, zd, r4 = O.error_indicators(m, PARAMS, x, x.copy(), 0.1)
    mu, lm = O.lame(PARAMS["E"], PARAMS["nu"])
    h, vol = _h_vol(m)
    want = h ** 2 * vol * (2 * (2 * mu + lm)) ** 2
    assert np.abs(zu - want).max() < 1e-10 * want.max()
    assert np.abs(zd).max() < 1e-12 * want.max() and np.abs(zp).max() < 1e-12 * want.max()
    assert np.abs(r4).max() == 0.0


def test_kink_across_a_facet():
    space, m = _two_cells()
    X = m["coords"]
    n = np.cross(X[2] - X[1], X[3] - X[1])
    n /= np.linalg.norm(n)
    if np.dot(X[0] - X[1], n) < 0:
        n = -n                                   # n points into cell 0
    e = np.array([0.3, -0.5, 0.8])
    e /= np.linalg.norm(e)
    s2 = np.maximum(0.0, (m["x2"] - X[1]) @ n)
    s1 = np.maximum(0.0, (X - X[1]) @ n)
    x = np.zeros(space.ndofs)
    x[: space.nu] = (s2[:, None] * e[None, :]).reshape(-1)
    x[space.nu::2] = s1                          # network 1
    zp, zu, zd, r4 = O.error_indicators(m, PARAMS, x, x.copy(), 0.1)
    mu, lm = O.lame(PARAMS["E"], PARAMS["nu"])
    h, vol = _h_vol(m)
    area = 0.5 * np.linalg.norm(np.cross(X[2] - X[1], X[3] - X[1]))
    havg = 0.5 * (h[0] + h[1])
    RE = mu * e + (mu + lm) * np.dot(e, n) * n
    Fu = havg * area * np.dot(RE, RE)
    Fp = havg * area * PARAMS["K"][0] ** 2
    # cell 0 also has RK_u = -alpha_1 grad p_1 = -alpha_1 n (the kink of p)
    cell0 = h[0] ** 2 * vol[0] * PARAMS["alpha"][0] ** 2
    assert np.allclose(zu, [Fu / 2 + cell0, Fu / 2], rtol=1e-12, atol=0)
    assert np.allclose(zp, [Fp / 2, Fp / 2], rtol=1e-12, atol=0)
    assert np.abs(zd).max() == 0.0 and np.abs(r4).max() == 0.0


def test_linear_pressure_time_and_exchange_terms():
    space, m = _box()
    s = 0.15
    params = dict(PARAMS, S=((0.0, s), (s, 0.0)))
    a = np.array([1.0, -0.5, 0.25])
    X = m["coords"]
    x = np.zeros(space.ndofs)
    x[space.nu::2] = X @ a                       # p_1 = a . X, p_2 = 0; u = 0
    dt = 0.05
    zp, zu, zd, r4 = O.error_indicators(m, params, x, np.zeros_like(x), dt)
    h, vol = _h_vol(m)
    v = X[m["cells"]]
    y = v @ a                                     # (nC, 4) nodal values of p_1
    # int_K p_1^2 = |K| / 20 (sum_k y_k^2 + (sum_k y_k)^2)
    ip2 = vol / 20 * ((y ** 2).sum(1) + y.sum(1) ** 2)
    c1, al1, K1 = params["c"][0], params["alpha"][0], params["K"][0]
    assert np.allclose(zp, h ** 2 * ((c1 / dt + s) ** 2 + s ** 2) * ip2, rtol=1e-11, atol=0)
    assert np.allclose(zu, h ** 2 * vol * al1 ** 2 * np.dot(a, a), rtol=1e-11, atol=0)
    assert np.allclose(zd, h ** 2 * vol * al1 ** 2 * np.dot(a, a) / dt ** 2, rtol=1e-11, atol=0)
    assert np.allclose(r4, K1 * np.dot(a, a) * vol + s * ip2, rtol=1e-11, atol=0)
