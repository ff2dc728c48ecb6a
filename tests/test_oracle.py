"""CPU tests pinning the FP64 oracle (and the product's reference tables).

Checks are chosen so that a plausible slip (dropped term, wrong sign,
transposed operand, wrong index) fails one of them:
  * element tensors == sympy-exact integration on 4 tetrahedra (fixtures);
  * the product's barycentric tables reproduce the same fixtures;
  * kernels: rigid motions annihilate eps:eps, constants the Laplacians;
    div of a linear field against P1 tests equals its integral;
  * symmetric elimination == row replacement (same solution);
  * known answer: u = 0, p_j = P_j, S = 0 is a steady discrete solution with
    p_0 = -sum_j alpha_j P_j;
  * textbook MINRES == scipy's MINRES;
  * 3D manufactured solution reaches the reference's convergence-rate
    thresholds (test_convergence_totalpressuresolver.py:237-242)."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import CFG1, oracle_mesh
from oracle import mpet_oracle as O
from oracle import mms

GOLD = os.path.join(os.path.dirname(__file__), "golden", "element_tensors.npz")


def _gold():
    return np.load(GOLD)


@pytest.mark.parametrize("k", range(4))
def test_oracle_element_tensors_match_sympy(k):
    d = _gold()
    et = O.element_tensors(d["coords"][k], np.array([[0, 1, 2, 3]]))
    for nm in "EGBMS":
        ref = d[f"{nm}{k}"]
        assert np.abs(et[nm][0] - ref).max() <= 1e-13 * np.abs(ref).max(), nm


@pytest.mark.parametrize("k", range(4))
def test_product_tables_match_sympy(k):
    """The GPU kernel combines W[a,b,k,l] g_k g_l^T and V[b,c,k] g_k; the
    same combination in numpy must give the exact tensors."""
    from waterscapes_amd import refel
    d = _gold()
    X = d["coords"][k]
    J = np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]], axis=1)
    vol = abs(np.linalg.det(J)) / 6
    Gi = np.linalg.inv(J)
    g = np.vstack([-Gi.sum(0), Gi])                       # grad lambda_k, rows
    W = refel.grad_weights_p2p2()
    V = refel.div_weights_p2p1()
    M = vol * np.einsum("abkl,km,ln->abmn", W, g, g)      # int d_m phi_a d_n phi_b
    G = np.einsum("abmm->ab", M)
    assert np.abs(G - d[f"G{k}"]).max() <= 1e-13 * np.abs(d[f"G{k}"]).max()
    E = np.zeros((10, 3, 10, 3))
    for i in range(3):
        for j in range(3):
            E[:, i, :, j] = 0.5 * ((i == j) * G + M[:, :, j, i])
    assert np.abs(E.reshape(30, 30) - d[f"E{k}"]).max() <= 1e-13 * np.abs(d[f"E{k}"]).max()
    B = vol * np.einsum("bck,km->bmc", V, g).reshape(30, 4)
    assert np.abs(B - d[f"B{k}"]).max() <= 1e-13 * np.abs(d[f"B{k}"]).max()
    Q = vol * refel.source_weights(1, 3)
    assert np.abs(Q - d[f"Q3_{k}"]).max() <= 1e-13 * np.abs(d[f"Q3_{k}"]).max()
    M1 = vol * refel.mass_p1()
    assert np.abs(M1 - d[f"M{k}"]).max() <= 1e-14 * np.abs(d[f"M{k}"]).max()


def test_element_kernels_and_closed_forms():
    rng = np.random.default_rng(0)
    X = rng.random((4, 3)) + np.eye(4, 3)
    et = O.element_tensors(X, np.array([[0, 1, 2, 3]]))
    # P2 node coordinates (vertices then UFC edges)
    nodes = np.vstack([X] + [0.5 * (X[i] + X[j]) for (i, j) in O.EDGES])
    modes = []
    for k in range(3):
        t = np.zeros((10, 3))
        t[:, k] = 1
        modes.append(t.reshape(-1))
    for (a, b) in ((0, 1), (0, 2), (1, 2)):
        r = np.zeros((10, 3))
        r[:, a] = -nodes[:, b]
        r[:, b] = nodes[:, a]
        modes.append(r.reshape(-1))
    scale = np.abs(et["E"][0]).max()
    for mvec in modes:
        assert np.abs(et["E"][0] @ mvec).max() <= 1e-13 * scale * 10
    assert np.abs(et["G"][0] @ np.ones(10)).max() <= 1e-13 * np.abs(et["G"][0]).max() * 10
    assert np.abs(et["S"][0] @ np.ones(4)).max() <= 1e-13 * np.abs(et["S"][0]).max() * 10
    vol = et["det"][0] / 6
    assert np.allclose(et["M"][0], vol / 20 * (np.ones((4, 4)) + np.eye(4)), rtol=1e-14)
    # u = x e_0 (div u = 1):  int psi_c div u = |K| / 4
    u = np.zeros((10, 3))
    u[:, 0] = nodes[:, 0]
    assert np.allclose(et["B"][0].T @ u.reshape(-1), vol / 4, rtol=1e-13)
    # cross-check P1 stiffness closed form |K| grad l_i . grad l_j
    J = np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]], axis=1)
    Gi = np.linalg.inv(J)
    g = np.vstack([-Gi.sum(0), Gi])
    assert np.allclose(et["S"][0], vol * g @ g.T, rtol=1e-13)


def _small_problem(n=2, A=2, params=None, **kw):
    from waterscapes_amd.fem import MixedSpace
    from waterscapes_amd.mesh import UnitCubeMesh
    mesh = UnitCubeMesh(n, n, n)
    space = MixedSpace(mesh, A)
    return mesh, space, oracle_mesh(space)


def test_symmetric_elimination_equals_row_replacement():
    mesh, space, m = _small_problem(2, 2)
    A = O.assemble_matrix(m["coords"], m["cells"], m["cell_p2"], m["cell_p1"], m["nP2"],
                          m["nP1"], CFG1, 0.01, 1.0, which="A")
    rng = np.random.default_rng(1)
    b = rng.standard_normal(A.shape[0])
    dofs = np.sort(rng.choice(A.shape[0], 40, replace=False))
    b[dofs] = rng.standard_normal(dofs.size)
    x_row = spla.spsolve(O.apply_row_replace(A, dofs).tocsc(), b.copy())
    bs = b.copy()
    As = O.apply_symmetric(A, dofs, bs)
    assert abs(As - As.T).max() <= 1e-12 * abs(A).max()
    x_sym = spla.spsolve(As.tocsc(), bs)
    assert np.allclose(x_row, x_sym, rtol=1e-9, atol=1e-12)
    assert np.allclose(x_sym[dofs], b[dofs])


def _all_dirichlet_markers(mesh):
    mk = np.full(mesh.num_facets(), 2**63 - 1, dtype=np.uint64)
    mk[mesh.exterior] = 0
    return mk


def test_uniform_pressure_known_answer():
    """u = 0, p_j = P_j constant, S = 0: a steady discrete solution with
    p_0 = -sum_j alpha_j P_j (the p_0 equation, :269)."""
    params = dict(CFG1)
    params["S"] = ((0.0, 0.0), (0.0, 0.0))
    P = (3.0, -1.5)
    mesh, space, m = _small_problem(2, 2)
    mk = _all_dirichlet_markers(mesh)
    lay = O.Layout(m["nP2"], m["nP1"], 2)
    up0 = np.zeros(lay.n)
    p0 = -(params["alpha"][0] * P[0] + params["alpha"][1] * P[1])
    up0[lay.pdofs(np.arange(m["nP1"]), 0)] = p0
    for j in range(2):
        up0[lay.pdofs(np.arange(m["nP1"]), j + 1)] = P[j]
    pb = [lambda X, t, v=P[j]: np.full(X.shape[0], v) for j in range(2)]
    prob = O.OracleProblem(m, params, dict(dt=0.01, theta=1.0, T=0.02),
                           u_bar=lambda X, t: np.zeros((X.shape[0], 3)), p_bar=pb,
                           momentum_markers=mk, continuity_markers=[mk, mk])
    for up, t in prob.solve_direct(up0):
        assert np.abs(up - up0).max() <= 1e-11 * np.abs(up0).max()
    for up, t in prob.solve_iterative(up0, rtol=1e-12):
        assert np.abs(up - up0).max() <= 1e-9 * np.abs(up0).max()


def test_textbook_minres_matches_scipy():
    mesh, space, m = _small_problem(2, 2)
    A = O.assemble_matrix(m["coords"], m["cells"], m["cell_p2"], m["cell_p1"], m["nP2"],
                          m["nP1"], CFG1, 0.01, 1.0, which="A")
    P = O.assemble_matrix(m["coords"], m["cells"], m["cell_p2"], m["cell_p1"], m["nP2"],
                          m["nP1"], CFG1, 0.01, 1.0, which="P")
    # make the un-eliminated P SPD with an identity shift on its u block
    P = P + sp.diags(np.r_[np.full(3 * m["nP2"], 1e2), np.zeros(A.shape[0] - 3 * m["nP2"])])
    rng = np.random.default_rng(5)
    b = rng.standard_normal(A.shape[0])
    lu = spla.splu(P.tocsc())
    A = A + sp.diags(np.r_[np.full(3 * m["nP2"], 1e2), np.zeros(A.shape[0] - 3 * m["nP2"])])
    x, it, _ = O.minres(lambda v: A @ v, b, lu.solve, rtol=1e-12)
    M = spla.LinearOperator(A.shape, matvec=lu.solve)
    xs, info = spla.minres(A, b, M=M, rtol=1e-12, maxiter=5000)
    xd = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xd) <= 1e-8 * np.linalg.norm(xd)
    # scipy stops on its own (looser) residual estimate
    assert np.linalg.norm(x - xs) <= 1e-4 * np.linalg.norm(xd)


def mms_run(n, params=CFG1, dt=0.05, T=0.1, theta=1.0):
    mesh, space, m = _small_problem(n, int(params["J"]))
    u, p, f, g = mms.exact_solutions(params)
    A = int(params["J"])
    mk = _all_dirichlet_markers(mesh)
    prob = O.OracleProblem(
        m, params, dict(dt=dt, theta=theta, T=T), f=mms.lambdify_vec(f),
        g=[mms.lambdify(gi) for gi in g], u_bar=mms.lambdify_vec(u),
        p_bar=[mms.lambdify(p[j + 1]) for j in range(A)], f_degree=3, g_degree=3,
        momentum_markers=mk, continuity_markers=[mk] * A)
    lay = prob.lay
    up0 = np.zeros(lay.n)
    uv = mms.lambdify_vec(u)(m["x2"], 0.0)
    up0[: lay.nu] = uv.reshape(-1)
    for j in range(A + 1):
        up0[lay.pdofs(np.arange(m["nP1"]), j)] = mms.lambdify(p[j])(m["coords"], 0.0)
    for up, t in prob.solve_direct(up0):
        pass
    errs = {}
    uh = up[: lay.nu].reshape(-1, 3)
    eu2 = eh1 = 0.0
    for k in range(3):
        l2, h1 = mms.error_norms(m["coords"], m["cells"], m["cell_p2"], m["x2"], uh[:, k],
                                 u[k], t, 2)
        eu2 += l2 ** 2
        eh1 += h1 ** 2
    errs["u"] = (np.sqrt(eu2), np.sqrt(eh1))
    for j in range(A + 1):
        vals = up[lay.pdofs(np.arange(m["nP1"]), j)]
        errs[f"p{j}"] = mms.error_norms(m["coords"], m["cells"], m["cell_p1"], m["coords"],
                                        vals, p[j], t, 1)
    return errs, up, t


def test_mms_convergence_rates_3d():
    """Rates between n = 4 and n = 8 meet the reference thresholds
    (u L2/H1 > 1.70, p_1 L2 > 1.85, p_2 L2 > 1.87, p H1 > 0.95)."""
    e4, _, _ = mms_run(4)
    e8, _, _ = mms_run(8)
    rate = lambda a, b: np.log(a / b) / np.log(2.0)  # noqa: E731
    ru2, ruh = rate(e4["u"][0], e8["u"][0]), rate(e4["u"][1], e8["u"][1])
    rp1, rp2 = rate(e4["p1"][0], e8["p1"][0]), rate(e4["p2"][0], e8["p2"][0])
    rh1, rh2 = rate(e4["p1"][1], e8["p1"][1]), rate(e4["p2"][1], e8["p2"][1])
    assert ru2 > 1.70 and ruh > 1.70, (ru2, ruh)
    assert rp1 > 1.85 and rp2 > 1.87, (rp1, rp2)
    assert rh1 > 0.95 and rh2 > 0.95, (rh1, rh2)


def test_library_exports_every_header_symbol():
    import ctypes
    from waterscapes_amd import _lib
    if not os.path.exists(_lib.LIBPATH):
        pytest.skip("library not built")
    L = ctypes.CDLL(_lib.LIBPATH)
    names = _lib.exported_symbols()
    assert len(names) >= 15
    for nm in names:
        assert hasattr(L, nm), nm


# ---------------------------------------------------------------------------
# exterior-facet terms (traction s, fluxes I_i, Robin beta_i / p_robin_i)
# ---------------------------------------------------------------------------
def _tri(seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((2, 3, 3)), np.array([[0, 1, 2], [3, 4, 5]])


def test_facet_integrals_closed_forms():
    """P1 facet mass |F|/12 (1 + delta); partition of unity of the P1 and
    P2 facet test functions; products of interpolants are exact."""
    X, tris = _tri(3)
    coords = X.reshape(-1, 3)
    v = coords[tris]
    area = 0.5 * np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), axis=1)
    M = O.facet_p1_mass(coords, tris)
    for f in range(2):
        assert np.allclose(M[f], area[f] / 12.0 * (np.ones((3, 3)) + np.eye(3)), rtol=1e-13)
    # beta constant: facet_source_matrix == beta * P1 facet mass
    Mb = O.facet_source_matrix(coords, tris, lambda P, t: np.full(P.shape[0], 2.5), 0, 0.0)
    assert np.allclose(Mb, 2.5 * M, rtol=1e-13)
    # linear data with P1 test: int f_h psi_a = M f_nodes (f linear is its own interpolant)
    lin = lambda P, t: (1.0 + P[:, 0] - 2.0 * P[:, 1] + 0.5 * P[:, 2]).reshape(-1, 1)  # noqa
    loc = O.facet_source_vector(coords, tris, 3, lin, 1, 0.0)[..., 0]
    fn = lin(coords, 0.0)[:, 0][tris]
    assert np.allclose(loc, np.einsum("fab,fb->fa", M, fn), rtol=1e-12)
    # partition of unity (P2 test functions): sum_a int f v_a = int f
    loc2 = O.facet_source_vector(coords, tris, 6, lin, 1, 0.0)[..., 0]
    assert np.allclose(loc2.sum(1), loc.sum(1), rtol=1e-12)
    # product of a constant and a linear interpolant
    pv = O.facet_product_vector(coords, tris, lambda P, t: np.full(P.shape[0], 3.0), 0,
                                lambda P, t: lin(P, t)[:, 0], 1, 0.0)
    assert np.allclose(pv, 3.0 * np.einsum("fab,fb->fa", M, fn), rtol=1e-12)


UNIT = dict(J=2, c=(1.0, 1.0), alpha=(1.0, 1.0), K=(1.0, 1.0), S=((1.0, 1.0), (1.0, 1.0)),
            E=1.0, nu=0.35)       # the reference test's parameters (:117-125)


def _right_face_markers(mesh, A):
    mid = mesh.facet_midpoints()
    right = mesh.exterior & (np.abs(mid[:, 0] - 1.0) < 1e-12)
    mk_u = _all_dirichlet_markers(mesh)
    mk_u[right] = 1
    mk_p = [_all_dirichlet_markers(mesh) for _ in range(A)]
    mk_p[0][right] = 1
    mk_p[1][right] = 2
    return mk_u, mk_p


def _facet_mms_problem(n, params, theta=1.0, dt=1.0, T=1.0):
    """Dirichlet everywhere except the face x = 1: traction s there for the
    momentum equation, flux I_1 for network 1 and a Robin condition for
    network 2 (beta = 2), all from the exact solution."""
    mesh, space, m = _small_problem(n, int(params["J"]))
    A = int(params["J"])
    # pressures with a non-zero normal derivative on x = 1
    pb = mms.S.cos(mms.S.pi * mms.X / 3) * mms.S.cos(mms.S.pi * mms.Y) * mms.S.cos(mms.S.pi * mms.Z)
    u, p, f, g = mms.exact_solutions(params, p_exprs=[(j + 1) * mms.T * pb for j in range(A)])
    sigma, grads = mms.boundary_data(params, u, p, 2.0)
    K = params["K"]
    mk_u, mk_p = _right_face_markers(mesh, A)
    sig = [[mms.lambdify(sigma[i][j]) for j in range(3)] for i in range(3)]

    def traction(P, t, nrm):
        S_ = np.stack([np.stack([sig[i][j](P, t) for j in range(3)], -1) for i in range(3)], -2)
        return np.einsum("nij,nj->ni", S_, nrm)
    # outward normal of the face x = 1 is e_x
    I1 = mms.lambdify(-K[0] * grads[0][0])
    pr2 = mms.lambdify(p[2] + K[1] * grads[1][0] / 2.0)
    prob = O.OracleProblem(
        m, params, dict(dt=dt, theta=theta, T=T), f=mms.lambdify_vec(f),
        g=[mms.lambdify(gi) for gi in g], u_bar=mms.lambdify_vec(u),
        p_bar=[mms.lambdify(p[j + 1]) for j in range(A)], f_degree=3, g_degree=3,
        momentum_markers=mk_u, continuity_markers=mk_p, s=traction, s_degree=3,
        I=[I1, None], I_degree=3, beta=[None, lambda P, t: np.full(P.shape[0], 2.0)],
        p_robin=[None, pr2], robin_degree=0, p_robin_degree=3)
    return mesh, space, m, prob, (u, p, f, g)


def _initial(prob, m, u, p):
    lay = prob.lay
    up0 = np.zeros(lay.n)
    up0[: lay.nu] = mms.lambdify_vec(u)(m["x2"], 0.0).reshape(-1)
    for j in range(len(p)):
        up0[lay.pdofs(np.arange(m["nP1"]), j)] = mms.lambdify(p[j])(m["coords"], 0.0)
    return up0


@pytest.mark.parametrize("theta", [1.0, 0.5])
def test_facet_terms_known_answer(theta):
    """Steady linear pressures with u = 0 (S = 0, so g = 0 and f = -grad p_0
    is constant) are represented exactly by P1/P2: with the exact traction
    (p_0 n), flux (-K grad p_1 . n) and Robin data on x = 1 and Dirichlet
    data elsewhere, one step must reproduce them to round-off.  A wrong sign,
    scale or time level of any facet term leaves an O(1) error."""
    params = dict(UNIT)
    params["S"] = ((0.0, 0.0), (0.0, 0.0))
    mesh, space, m = _small_problem(2, 2)
    Xs, Ys, Zs = mms.X, mms.Y, mms.Z
    pj = [1.0 + 0.5 * Xs + 0.2 * Ys, 2.0 - 0.3 * Xs + 0.1 * Zs]
    p0 = -(pj[0] + pj[1])
    u = [0 * Xs, 0 * Xs, 0 * Xs]
    f = [-mms.S.diff(p0, v) for v in (Xs, Ys, Zs)]
    mk_u, mk_p = _right_face_markers(mesh, 2)
    P0 = mms.lambdify(p0)
    prob = O.OracleProblem(
        m, params, dict(dt=0.7, theta=theta, T=0.7), f=mms.lambdify_vec(f),
        g=None, u_bar=mms.lambdify_vec(u), p_bar=[mms.lambdify(q) for q in pj], f_degree=1,
        momentum_markers=mk_u, continuity_markers=mk_p,
        s=lambda P, t, nrm: P0(P, t)[:, None] * nrm, s_degree=1,
        I=[mms.lambdify(-1.0 * mms.S.diff(pj[0], Xs)), None], I_degree=0,
        beta=[None, lambda P, t: np.full(P.shape[0], 2.0)],
        p_robin=[None, mms.lambdify(pj[1] + mms.S.diff(pj[1], Xs) / 2.0)], robin_degree=0,
        p_robin_degree=1)
    exact = _initial(prob, m, u, [p0] + pj)
    (up, t), = list(prob.solve_direct(exact.copy()))
    assert np.abs(up - exact).max() <= 1e-10 * np.abs(exact).max(), np.abs(up - exact).max()
    # a flipped flux sign is detected
    b_ok, _, _ = prob.rhs(exact, 0.0)
    prob.I = [mms.lambdify(1.0 * mms.S.diff(pj[0], Xs)), None]
    b_bad, _, _ = prob.rhs(exact, 0.0)
    assert np.abs(b_bad - b_ok).max() > 1e-3


def test_mms_convergence_rates_with_facet_terms():
    """Traction, flux and Robin data derived from an exact solution with
    non-zero normal derivatives on x = 1 keep second-order L2 rates."""
    rate = lambda a, b: np.log(a / b) / np.log(2.0)  # noqa: E731
    errs = []
    for n in (3, 6):
        mesh, space, m, prob, (u, p, f, g) = _facet_mms_problem(n, UNIT)
        for up, t in prob.solve_direct(_initial(prob, m, u, p)):
            pass
        lay = prob.lay
        e = {}
        uh = up[: lay.nu].reshape(-1, 3)
        e["u"] = np.sqrt(sum(mms.error_norms(m["coords"], m["cells"], m["cell_p2"], m["x2"],
                                             uh[:, k], u[k], t, 2)[0] ** 2 for k in range(3)))
        for j in (1, 2):
            e[f"p{j}"] = mms.error_norms(m["coords"], m["cells"], m["cell_p1"], m["coords"],
                                         up[lay.pdofs(np.arange(m["nP1"]), j)], p[j], t, 1)[0]
        errs.append(e)
    ru = rate(errs[0]["u"], errs[1]["u"])
    rp1, rp2 = rate(errs[0]["p1"], errs[1]["p1"]), rate(errs[0]["p2"], errs[1]["p2"])
    assert ru > 1.70 and rp1 > 1.70 and rp2 > 1.70, (ru, rp1, rp2, errs)


def _bench_like(n, A, steps, scale_K=1.0, seed=7):
    """bench.py's workload family at n^3 (u = 0 on z = 0, p_j = P0_j +
    5 sin(2 pi t) on z = 1, log-normal per-cell K, adjacent exchange)."""
    from conftest import cfg2_params, lognormal_K
    mesh, space, m = _small_problem(n, A)
    params = cfg2_params(A)
    Kc = lognormal_K(mesh.num_cells(), params["K"], seed=seed) * scale_K
    mid = mesh.facet_midpoints()
    inv = np.uint64(2**63 - 1)
    mk_u = np.full(mesh.num_facets(), inv, dtype=np.uint64)
    mk_u[mesh.exterior & (np.abs(mid[:, 2]) < 1e-12)] = 0
    mk_p = np.full(mesh.num_facets(), inv, dtype=np.uint64)
    mk_p[mesh.exterior & (np.abs(mid[:, 2] - 1.0) < 1e-12)] = 0
    P0 = (8.0, 4.0, 2.0, 1.0)
    pb = [lambda X, t, j=j: np.full(X.shape[0], P0[j] + 5.0 * np.sin(2.0 * np.pi * t))
          for j in range(A)]
    prob = O.OracleProblem(m, params, dict(dt=0.0125, theta=1.0, T=steps * 0.0125),
                           u_bar=lambda X, t: np.zeros((X.shape[0], 3)), p_bar=pb, K_cells=Kc,
                           momentum_markers=mk_u, continuity_markers=[mk_p] * A)
    return m, params, Kc, prob


def test_sampled_row_residuals_pin_the_full_assembly():
    """The oracle's equations at a sample of rows, assembled from only the
    cells that touch them (the check the 48^3 GPU step runs at full size):
    the rows equal the full assembly's, the oracle's direct solution
    satisfies them to round-off, and a 1% error in K or in E shows."""
    from conftest import sampled_row_residuals, field_row_ratios
    A = 2
    m, params, Kc, prob = _bench_like(4, A, 2)
    lay = prob.lay
    x0 = np.zeros(lay.n)
    states = [x.copy() for x, t in prob.solve_direct(x0)]
    x_prev, x = states[0], states[1]
    bc = np.unique(np.concatenate(prob.bc_dofs))
    free = np.setdiff1d(np.arange(lay.n), bc)
    rng = np.random.default_rng(3)
    rows = np.sort(rng.choice(free, 120, replace=False))
    # restricted rows == full rows
    r, scale = sampled_row_residuals(m, params, Kc, 0.0125, 1.0, x, x_prev, rows)
    full = prob.A_mat[rows] @ x - prob.B_matrix()[rows] @ x_prev
    assert np.allclose(r, full, rtol=0, atol=1e-13 * np.max(scale))
    ratios = field_row_ratios(lay, rows, r, scale)
    assert len(ratios) == A + 2 and max(ratios) <= 1e-12, ratios
    # planted slips in the assembly of the check: 1% in K, 1% in E
    r1, s1 = sampled_row_residuals(m, params, Kc * 1.01, 0.0125, 1.0, x, x_prev, rows)
    assert max(field_row_ratios(lay, rows, r1, s1)[2:]) >= 1e-6
    pe = dict(params, E=params["E"] * 1.01)
    r2, s2 = sampled_row_residuals(m, pe, Kc, 0.0125, 1.0, x, x_prev, rows)
    assert field_row_ratios(lay, rows, r2, s2)[0] >= 1e-4
