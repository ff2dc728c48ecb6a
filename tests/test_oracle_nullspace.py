"""Oracle pinning of the nullspace multipliers (rigid motions of u, constants
of chosen pressures: mpet/rm_basis_L2.py:10-74, mpettotalpressuresolver.py:
179-202, 285-306, mpetsolver.py:107-128, 203-215).

* the six rigid motions are L2-orthonormal (quadrature Gram matrix = I) and
  match the product's host construction (waterscapes_amd/nullspace.py, closed
  forms instead of quadrature) up to the sign of each basis function;
* the constraint rows reproduce the Gram matrix on the interpolated rigid
  motions (B Zt = I) and the volume on constants;
* the rigid motions span a kernel of A (both formulations), which is what
  makes the multiplier rows necessary;
* known answers: uniform pressures with the matching traction on every
  facet give u = 0 (total pressure: p_0 = -sum alpha_j P_j; two-field: the
  3D version of test_donut.py:16-71); a uniform dilation load with
  pure-Neumann pressures (test_donut.py:73-131 in 3D) gives
  u = t/(2 mu + 3 lambda) (x - c), p_j = 0 and pressure multipliers
  3 alpha_j (gamma_{n+1} - gamma_n)."""
import numpy as np
import pytest

from conftest import oracle_mesh
from oracle import mpet_oracle as O


def _mesh(n=(3, 4, 2), total_pressure=True, A=2):
    from waterscapes_amd.fem import MixedSpace
    from waterscapes_amd.mesh import BoxMesh
    mesh = BoxMesh((0.0, 0.0, 0.0), (1.0, 1.5, 0.7), *n)
    space = MixedSpace(mesh, A, total_pressure=total_pressure)
    return space, oracle_mesh(space)


def _gram(m, Z):
    J, det, _ = O.cell_geometry(m["coords"], m["cells"])
    qp, qw = O.tet_rule(2)
    X = m["coords"][m["cells"][:, 0]][:, None, :] + np.einsum("cij,qj->cqi", J, qp)
    w = det[:, None] * qw[None, :]
    V = [np.asarray(z(X.reshape(-1, 3))).reshape(X.shape) for z in Z]
    return np.array([[np.einsum("cq,cqk,cqk->", w, a, b) for b in V] for a in V])


def test_rigid_motions_orthonormal_and_match_product():
    from waterscapes_amd.nullspace import RigidMotions
    space, m = _mesh()
    Z, vol = O.rigid_motions(m["coords"], m["cells"])
    assert abs(vol - 1.0 * 1.5 * 0.7) < 1e-13
    G = _gram(m, Z)
    assert np.abs(G - np.eye(6)).max() < 1e-12
    rm = RigidMotions(m["coords"], m["cells"])
    X = m["x2"]
    got = rm.values(X)
    for k in range(6):
        ref = Z[k](X)
        s = np.sign(np.vdot(got[k], ref))
        assert np.abs(s * got[k] - ref).max() < 1e-12 * np.abs(ref).max()
    assert abs(rm.volume - vol) < 1e-13


def test_constraint_rows_reproduce_gram_and_volume():
    space, m = _mesh()
    lay = O.Layout(m["nP2"], m["nP1"], 2)
    Z, vol = O.rigid_motions(m["coords"], m["cells"])
    C = O.constraint_rows(m["coords"], m["cells"], m["cell_p2"], m["cell_p1"], lay, Z, [0, 2])
    assert C.shape == (8, lay.n)
    Zt = np.zeros((lay.n, 6))
    for k in range(6):
        Zt[: lay.nu, k] = Z[k](m["x2"]).reshape(-1)
    assert np.abs(C[:6] @ Zt - np.eye(6)).max() < 1e-12
    for j, f in enumerate((0, 2)):
        one = np.zeros(lay.n)
        one[lay.nu + f::lay.F] = 1.0
        assert abs(C[6 + j] @ one - vol) < 1e-12
        other = np.zeros(lay.n)
        other[lay.nu + 1::lay.F] = 1.0
        assert C[6 + j] @ other == 0.0
        assert np.abs(C[6 + j, : lay.nu].toarray()).max() == 0.0


@pytest.mark.parametrize("formulation", ["total_pressure", "two_field"])
def test_rigid_motions_span_a_kernel_of_A(formulation):
    tp = formulation == "total_pressure"
    space, m = _mesh(total_pressure=tp)
    params = dict(J=2, E=500.0, nu=0.35, alpha=(1.0, 0.5), c=(1e-3, 0.0), K=(1e-2, 1e-1),
                  S=((0.0, 1e-2), (1e-2, 0.0)))
    prob = O.OracleProblem(m, params, dict(dt=0.1, theta=1.0, T=0.1), formulation=formulation)
    Z, _ = O.rigid_motions(m["coords"], m["cells"])
    lay = prob.lay
    for k in range(6):
        x = np.zeros(lay.n)
        x[: lay.nu] = Z[k](m["x2"]).reshape(-1)
        r = prob.A_mat @ x
        assert np.abs(r).max() < 1e-10 * abs(prob.A_mat).max() * np.abs(x).max()


def _uniform_pressure_problem(formulation, A=2):
    """u nullspace, traction sigma n = -sum alpha_j P(t) n on every facet,
    p_j = P(t) = -t Dirichlet on every facet, c = S = 0."""
    tp = formulation == "total_pressure"
    space, m = _mesh((2, 3, 2), total_pressure=tp, A=A)
    alpha = (1.0, 0.5)[:A]
    params = dict(J=A, E=500.0, nu=0.35, alpha=alpha, c=(0.0,) * A, K=(1e-2, 1e-1)[:A],
                  S=[[0.0] * A for _ in range(A)])
    sa = sum(alpha)
    nF = m["facets"].shape[0]
    mm = np.full(nF, 1, dtype=np.uint64)                   # traction everywhere
    cm = [np.zeros(nF, dtype=np.uint64) for _ in range(A)]  # Dirichlet everywhere
    prob = O.OracleProblem(
        m, params, dict(dt=0.1, theta=1.0, T=0.2),
        u_bar=lambda X, t: np.zeros((X.shape[0], 3)),
        p_bar=[lambda X, t: np.full(X.shape[0], -t) for _ in range(A)],
        s=lambda X, t, nrm: sa * t * nrm, s_degree=1,
        momentum_markers=mm, continuity_markers=cm, formulation=formulation, u_nullspace=True)
    return prob, params


@pytest.mark.parametrize("formulation", ["total_pressure", "two_field"])
def test_uniform_pressure_known_answer(formulation):
    prob, params = _uniform_pressure_problem(formulation)
    lay = prob.lay
    assert prob.nmult == 6
    steps = list(prob.solve_direct(np.zeros(lay.n + 6)))
    assert len(steps) == 2
    sa = sum(params["alpha"])
    for x, t in steps:
        assert np.abs(x[: lay.nu]).max() < 1e-12
        for i in range(params["J"]):
            p = x[lay.nu + lay.net(i)::lay.F][: lay.nP1]
            assert np.abs(p + t).max() < 1e-12
        if formulation == "total_pressure":
            p0 = x[lay.nu: lay.n][0::lay.F]
            assert np.abs(p0 - sa * t).max() < 1e-12
        assert np.abs(x[lay.n:]).max() < 1e-12


def test_dilation_with_pressure_nullspaces_known_answer():
    """Two-field, u and both pressures with a nullspace (pure Neumann),
    traction t n on every facet, c = S = 0: u = gamma(t) (x - c),
    gamma = t / (2 mu + 3 lambda); p_j = 0; rigid-motion multipliers 0;
    pressure multipliers 3 alpha_j (gamma_{n+1} - gamma_n)."""
    space, m = _mesh((2, 3, 2), total_pressure=False)
    params = dict(J=2, E=500.0, nu=0.35, alpha=(1.0, 0.5), c=(0.0, 0.0), K=(1e-2, 1e-1),
                  S=((0.0, 0.0), (0.0, 0.0)))
    nF = m["facets"].shape[0]
    mm = np.full(nF, 1, dtype=np.uint64)
    cm = [np.full(nF, 1, dtype=np.uint64), np.full(nF, 2**63 - 1, dtype=np.uint64)]
    prob = O.OracleProblem(
        m, params, dict(dt=0.1, theta=1.0, T=0.2), s=lambda X, t, nrm: t * nrm, s_degree=1,
        momentum_markers=mm, continuity_markers=cm, formulation="two_field",
        u_nullspace=True, p_nullspace=(True, True))
    lay = prob.lay
    assert prob.nmult == 8
    mu, lmbda = O.lame(params["E"], params["nu"])
    _, vol = O.rigid_motions(m["coords"], m["cells"])
    cen = np.array([0.5, 0.75, 0.35])
    t_prev = 0.0
    for x, t in prob.solve_direct(np.zeros(lay.n + 8)):
        g = t / (2 * mu + 3 * lmbda)
        u = g * (m["x2"] - cen)
        assert np.abs(x[: lay.nu] - u.reshape(-1)).max() < 1e-12 * np.abs(u).max()
        assert np.abs(x[lay.nu: lay.n]).max() < 1e-10
        assert np.abs(x[lay.n: lay.n + 6]).max() < 1e-10
        dg = (t - t_prev) / (2 * mu + 3 * lmbda)
        want = 3 * np.asarray(params["alpha"]) * dg
        assert np.abs(x[lay.n + 6:] - want).max() < 1e-10 * np.abs(want).max()
        t_prev = t


def test_mixed_space_multiplier_layout():
    """Product host layout: the multipliers follow the field DOFs (u group of
    six, then one per constrained pressure), split() returns them after the
    fields, as the reference's up.split() returns its R components."""
    from waterscapes_amd.fem import Function, MixedSpace
    from waterscapes_amd.mesh import BoxMesh
    mesh = BoxMesh((0.0, 0.0, 0.0), (1.0, 1.5, 0.7), 2, 3, 2)
    sp_ = MixedSpace(mesh, 2, u_nullspace=True, p_fields=[0, 2])
    assert sp_.nmult == 8 and sp_.dim() == sp_.ndofs + 8
    assert sp_.num_sub_spaces() == sp_.F + 1 + 3
    f = Function(sp_)
    f.data[:] = np.arange(sp_.dim(), dtype=np.float64)
    parts = f.split()
    assert len(parts) == sp_.F + 4
    assert np.array_equal(parts[sp_.F + 1].vector().get_local(), np.arange(sp_.ndofs, sp_.ndofs + 6))
    assert parts[sp_.F + 3].vector().get_local().tolist() == [sp_.ndofs + 7]
    assert np.array_equal(parts[2].vector().get_local(),
                          sp_.nu + 1 + sp_.F * np.arange(sp_.nP1))
    with pytest.raises(ValueError):
        MixedSpace(mesh, 2, p_fields=[3])


def test_new_entry_points_reject_invalid_arguments():
    """Argument checks of the constraint-row, P2-mass, bordered-solver and
    indicator entry points return MPET_ERR_INVALID with a message before
    any device work (host-only calls)."""
    import ctypes as C
    import os
    from waterscapes_amd import _lib
    if not os.path.exists(_lib.LIBPATH):
        pytest.skip("library not built")
    L = _lib.lib()
    P = C.c_void_p
    L.mpet_solver_create_bordered.argtypes = [P, P, C.c_int, C.c_int, C.c_int, P, P, C.c_int,
                                              C.c_double, P, C.POINTER(C.c_void_p)]
    h = C.c_void_p()
    msg = C.create_string_buffer(256)

    def err(rc):
        assert rc == 1                       # MPET_ERR_INVALID
        L.mpet_last_error(msg, 256)
        return msg.value.decode()

    assert "at least one constraint" in err(
        L.mpet_solver_create_bordered(None, None, 1, 1, 1, None, None, 0, 1.0, None, C.byref(h)))
    assert "invalid argument" in err(
        L.mpet_solver_create_bordered(None, None, 1, 1, 1, None, None, 6, 1.0, None, C.byref(h)))
    assert "invalid argument" in err(
        L.mpet_constraint_rows(None, None, None, 1, 1, 1, None, 6, None, 0, None, None))
    assert "invalid argument" in err(L.mpet_assemble_p2_mass(None, None, None, None))
    assert "invalid argument" in err(
        L.mpet_error_indicators(None, None, None, None, 0, None, None, None, None))
    assert h.value is None


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6])
def test_extrapolated_guess_is_exact_for_polynomials(order):
    """Host logic of the initial guess (StepEngine, MPETBatch): with the
    history filled, the guess reproduces states that are polynomials of
    degree `order` in time exactly; before that it uses the states it has."""
    import torch
    from waterscapes_amd.solver import extrapolated_guess, push_history
    coef = torch.tensor([[0.3, -1.0, 2.0, 0.5, -0.25, 0.125, -0.0625][: order + 1]],
                        dtype=torch.float64)

    def state(t):
        return (coef * torch.tensor([t ** k for k in range(order + 1)],
                                    dtype=torch.float64)).sum(1).repeat(3)

    hist = []
    x = torch.empty(3, dtype=torch.float64)
    for n in range(order + 4):
        extrapolated_guess(hist, order, state(n), x)
        if len(hist) >= order:
            assert torch.allclose(x, state(n + 1), rtol=1e-12, atol=1e-12)
        elif not hist:
            assert torch.equal(x, state(n))
        push_history(hist, order, state(n))
        assert len(hist) == min(n + 1, order)
        assert torch.equal(hist[0], state(n))


def test_adaptive_order_picks_the_degree_that_predicted_best():
    """Host logic of the adaptive initial guess: on states that are a
    quadratic in time the degree picked reproduces the last state (error at
    rounding level); just after a kink in time (a start-up transient) the
    linear guess, which reaches back least, is picked over the high
    degrees."""
    import torch
    from waterscapes_amd.solver import adaptive_order, extrapolated_guess, push_history
    nu, F, nP1 = 6, 3, 4
    n = nu + F * nP1
    rng = torch.Generator().manual_seed(3)
    a, b, c = (torch.randn(n, generator=rng, dtype=torch.float64) for _ in range(3))

    def quad(t):
        return a + b * t + c * t * t
    hist, tmp, x = [], torch.empty(n, dtype=torch.float64), torch.empty(n, dtype=torch.float64)
    for t in range(8):
        push_history(hist, 7, quad(float(t)))
    cur = quad(8.0)
    k = adaptive_order(hist, 6, cur, nu, F, nP1, tmp)
    extrapolated_guess(hist[1:], k, hist[0], x)
    assert k >= 2 and torch.allclose(x, cur, rtol=1e-10, atol=1e-10)

    def kinked(t):
        return a * max(t - 5.0, 0.0) + b      # flat, then linear from t = 5
    hist = []
    for t in range(7):
        push_history(hist, 7, kinked(float(t)))
    assert adaptive_order(hist, 6, kinked(7.0), nu, F, nP1, tmp) == 1
