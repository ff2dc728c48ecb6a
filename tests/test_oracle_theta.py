"""CPU tests pinning the theta-weighted terms of the oracle (theta != 1).

The theta scheme of mpettotalpressuresolver.py:239-241, :270-274 weights the
network stiffness and the exchange S_ij at t_{n+1} by dt*theta (matrix A)
and at t_n by dt*(1 - theta) (the up_-linear part of b); sources and fluxes
are taken at t_theta = t_n + theta*dt (:467-476).  theta = 0.5 is the
reference default (:146-162) and the setting of the reference's own
Crank-Nicolson convergence experiment
(test_convergence_totalpressuresolver.py:103-125, :199-204, :247-249).

  * known answer: pressures linear in space AND time, u = 0, exchange
    between three networks, a flux face x = 1.  P1 represents such fields
    exactly and the theta scheme integrates linear-in-time solutions
    exactly when the data are taken at t_theta, so every step must
    reproduce the exact nodal values to round-off for every theta.  The
    expected values come from the PDE (sympy), not from the oracle's
    coefficients;
  * planted slips in the oracle (dt*S for dt*theta*S in A, a flipped or
    un-weighted (1 - theta) exchange in b, stiffness weights with theta and
    1 - theta swapped) break that known answer;
  * the reference's theta = 0.5 experiment in 3D (J = 2, c = alpha = K = 1,
    S = ((1,1),(1,1)), E = 1, nu = 0.35, traction face x = 1, paired mesh
    and time refinement) reaches the reference's rate thresholds."""
import numpy as np
import pytest

from oracle import mpet_oracle as O
from oracle import mms
from test_oracle import _all_dirichlet_markers, _initial, _small_problem

S_ = mms.S

# three networks with distinct exchange coefficients between every pair
KA_PARAMS = dict(J=3, E=10.0, nu=0.3, alpha=(0.5, 1.0, 0.25), c=(1.0, 0.5, 2.0),
                 K=(1.0, 2.0, 0.5), S=((0.0, 1.0, 2.0), (1.0, 0.0, 0.5), (2.0, 0.5, 0.0)))


def known_answer_fields(params=KA_PARAMS):
    """Exact fields (sympy): u = 0, p_j = P_j(x) + t Q_j(x) with P_j, Q_j
    linear, p_0 = -sum_j alpha_j p_j; f, g from the PDE with the reference's
    negated source convention, flux I_j = -K_j dp_j/dx on x = 1."""
    X, Y, Z, T = mms.X, mms.Y, mms.Z, mms.T
    A = int(params["J"])
    pj = [(1.0 + j) * (1.0 + 0.5 * X - 0.3 * Y + 0.2 * Z)
          + T * ((2.0 - j) - (0.7 + 0.4 * j) * X + 0.5 * Z + 0.3 * j * Y)
          for j in range(A)]
    u, p, f, g = mms.exact_solutions(params, u_expr=[0 * X, 0 * X, 0 * X], p_exprs=pj)
    flux = [-params["K"][j] * S_.diff(p[j + 1], X) for j in range(A)]
    return u, p, f, g, flux


def _known_answer_markers(mesh, A):
    mid = mesh.facet_midpoints()
    right = mesh.exterior & (np.abs(mid[:, 0] - 1.0) < 1e-12)
    mk_u = _all_dirichlet_markers(mesh)
    mk_p = [_all_dirichlet_markers(mesh) for _ in range(A)]
    for mk in mk_p:
        mk[right] = 1
    return mk_u, mk_p


def known_answer_problem(n, theta, dt, T, params=KA_PARAMS):
    mesh, space, m = _small_problem(n, int(params["J"]))
    A = int(params["J"])
    u, p, f, g, flux = known_answer_fields(params)
    mk_u, mk_p = _known_answer_markers(mesh, A)
    prob = O.OracleProblem(
        m, params, dict(dt=dt, theta=theta, T=T), f=mms.lambdify_vec(f),
        g=[mms.lambdify(gi) for gi in g], u_bar=mms.lambdify_vec(u),
        p_bar=[mms.lambdify(p[j + 1]) for j in range(A)], f_degree=1, g_degree=1,
        momentum_markers=mk_u, continuity_markers=mk_p,
        I=[mms.lambdify(q) for q in flux], I_degree=1)
    return mesh, space, m, prob, (u, p, f, g)


def _exact_at(prob, m, sym, t):
    u, p, f, g = sym
    lay = prob.lay
    x = np.zeros(lay.n)
    for j in range(len(p)):
        x[lay.pdofs(np.arange(m["nP1"]), j)] = mms.lambdify(p[j])(m["coords"], t)
    return x


def _known_answer_errors(theta, dt=0.25, T=0.75, n=2):
    mesh, space, m, prob, sym = known_answer_problem(n, theta, dt, T)
    up0 = _exact_at(prob, m, sym, 0.0)
    errs = []
    for up, t in prob.solve_direct(up0.copy()):
        ex = _exact_at(prob, m, sym, t)
        errs.append(np.abs(up - ex).max() / np.abs(ex).max())
    return errs


@pytest.mark.parametrize("theta", [0.5, 0.3, 1.0])
def test_theta_exchange_known_answer(theta):
    errs = _known_answer_errors(theta)
    assert len(errs) == 3
    assert max(errs) <= 1e-11, errs


def _exchange_parts(S, dt, A):
    """(i, j, w) with w the exchange weight per unit dt of entry (i+1, j+1)
    of the pressure coefficient matrices: -sum_{k != i} S_ik on the
    diagonal, S_ij off it."""
    out = []
    for i in range(A):
        for j in range(A):
            w = -sum(S[i][k] for k in range(A) if k != i) if i == j else S[i][j]
            out.append((i, j, w))
    return out


def _slip(kind):
    orig = O.pressure_coefficients

    def slipped(alpha, c, S, lmbda, dt, theta):
        Cm, Bm = orig(alpha, c, S, lmbda, dt, theta)
        A = len(alpha)
        for i, j, w in _exchange_parts(S, dt, A):
            if kind == "A_dt_not_dt_theta":          # lhs exchange with dt
                Cm[i + 1, j + 1] += dt * (1.0 - theta) * w
            elif kind == "b_sign":                   # (1 - theta) part flipped
                Bm[i + 1, j + 1] += 2.0 * dt * (1.0 - theta) * w
            elif kind == "b_dt_not_dt_1m_theta":     # rhs exchange with dt
                Bm[i + 1, j + 1] += -dt * theta * w
        return Cm, Bm
    return slipped


@pytest.mark.parametrize("kind", ["A_dt_not_dt_theta", "b_sign", "b_dt_not_dt_1m_theta"])
def test_theta_exchange_planted_slips_fail(kind, monkeypatch):
    monkeypatch.setattr(O, "pressure_coefficients", _slip(kind))
    errs = _known_answer_errors(0.5)
    assert max(errs) > 1e-4, (kind, errs)


def test_theta_stiffness_weights_swapped_fail():
    """The stiffness and exchange of b taken with 1 - theta in place of
    theta (A keeps theta): the known answer breaks at theta = 0.3."""
    mesh, space, m, prob, sym = known_answer_problem(2, 0.3, 0.25, 0.75)
    bad = known_answer_problem(2, 0.7, 0.25, 0.75)[3]
    prob._B_mat = bad.B_matrix()
    up0 = _exact_at(prob, m, sym, 0.0)
    (up, t), = [next(iter(prob.solve_direct(up0.copy())))]
    ex = _exact_at(prob, m, sym, t)
    assert np.abs(up - ex).max() / np.abs(ex).max() > 1e-4


UNIT = dict(J=2, c=(1.0, 1.0), alpha=(1.0, 1.0), K=(1.0, 1.0), S=((1.0, 1.0), (1.0, 1.0)),
            E=1.0, nu=0.35)       # test_convergence_totalpressuresolver.py:111-118


def reference_theta_fields(params=UNIT):
    """3D counterpart of the reference experiment's fields (:38-45):
    u_k = cos(2 pi x) cos(2 pi y) cos(2 pi z) sin((2 pi + 1) t) for each
    component, p_i = -i times the same; p_0 from the constraint."""
    X, Y, Z, T = mms.X, mms.Y, mms.Z, mms.T
    A = int(params["J"])
    sp = (S_.cos(2 * S_.pi * X) * S_.cos(2 * S_.pi * Y) * S_.cos(2 * S_.pi * Z)
          * S_.sin(2 * S_.pi * T + T))
    return mms.exact_solutions(params, u_expr=[sp, sp, sp],
                               p_exprs=[-(i + 1) * sp for i in range(A)])


def reference_theta_problem(n, M, theta=0.5, params=UNIT):
    """Dirichlet everywhere, except traction sigma.n on x = 1 for the
    momentum equation (:140-147); T = 1, dt = 1/M."""
    mesh, space, m = _small_problem(n, int(params["J"]))
    A = int(params["J"])
    u, p, f, g = reference_theta_fields(params)
    sigma, _ = mms.boundary_data(params, u, p, 1.0)
    sig = [[mms.lambdify(sigma[i][j]) for j in range(3)] for i in range(3)]

    def traction(P, t, nrm):
        St = np.stack([np.stack([sig[i][j](P, t) for j in range(3)], -1) for i in range(3)], -2)
        return np.einsum("nij,nj->ni", St, nrm)
    mid = mesh.facet_midpoints()
    right = mesh.exterior & (np.abs(mid[:, 0] - 1.0) < 1e-12)
    mk_u = _all_dirichlet_markers(mesh)
    mk_u[right] = 1
    mk_p = _all_dirichlet_markers(mesh)
    prob = O.OracleProblem(
        m, params, dict(dt=1.0 / M, theta=theta, T=1.0), f=mms.lambdify_vec(f),
        g=[mms.lambdify(gi) for gi in g], u_bar=mms.lambdify_vec(u),
        p_bar=[mms.lambdify(p[j + 1]) for j in range(A)], f_degree=3, g_degree=3,
        momentum_markers=mk_u, continuity_markers=[mk_p] * A, s=traction, s_degree=3)
    return mesh, space, m, prob, (u, p, f, g)


def _errors(prob, m, sym, up, t):
    u, p, f, g = sym
    lay = prob.lay
    uh = up[: lay.nu].reshape(-1, 3)
    l2 = h1 = 0.0
    for k in range(3):
        a, b = mms.error_norms(m["coords"], m["cells"], m["cell_p2"], m["x2"], uh[:, k], u[k],
                               t, 2)
        l2, h1 = l2 + a * a, h1 + b * b
    e = {"uL2": np.sqrt(l2), "uH1": np.sqrt(h1)}
    for j in (1, 2):
        e[f"p{j}L2"], e[f"p{j}H1"] = mms.error_norms(
            m["coords"], m["cells"], m["cell_p1"], m["coords"],
            up[lay.pdofs(np.arange(m["nP1"]), j)], p[j], t, 1)
    return e


def test_reference_theta_half_experiment_rates_3d():
    """(n, M) = (4, 2) -> (8, 4) (the reference pairs n = 8, 16, 32 with
    M = 4, 8, 16 in 2D): u L2/H1 > 1.70, p_1 L2 > 1.85, p_2 L2 > 1.87,
    p H1 > 0.95 (:237-242)."""
    errs = []
    for n, M in ((4, 2), (8, 4)):
        mesh, space, m, prob, sym = reference_theta_problem(n, M)
        for up, t in prob.solve_direct(_initial(prob, m, sym[0], sym[1])):
            pass
        assert abs(t - 1.0) < 1e-12
        errs.append(_errors(prob, m, sym, up, t))
    r = {k: np.log(errs[0][k] / errs[1][k]) / np.log(2.0) for k in errs[0]}
    assert r["uL2"] > 1.70 and r["uH1"] > 1.70, r
    assert r["p1L2"] > 1.85 and r["p2L2"] > 1.87, r
    assert r["p1H1"] > 0.95 and r["p2H1"] > 0.95, r
