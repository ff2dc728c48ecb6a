"""Two-field MPETSolver (mpet/mpetsolver.py) on the GPU vs the FP64 oracle.

* the device-expanded matrices of lhs(F) (with Robin), prec and the
  up_-linear part of rhs(F) equal the oracle's assembly (1e-12 relative);
* every step of solve() (the reference's direct path) and of step(dt)
  matches the oracle's row-replaced LU solve (:382-462) to 1e-8 relative L2
  per field: the cfg1 manufactured solution with Dirichlet data on every
  face, and a case with traction, flux and Robin facets (markers 1 and 2);
* the problem is set up like the reference's MPETSolver test
  (test_convergence_mpetsolver.py:120-160): Expressions, CompiledSubDomain
  markers, assign(up_.sub(i), interpolate(...)), then solver.solve()."""
import numpy as np
import pytest

from conftest import CFG1, oracle_mesh
from oracle import mpet_oracle as O
from oracle import mms
from test_gpu_api import cstr

pytestmark = pytest.mark.gpu


def build(n, params=CFG1, T=0.1, dt=0.05, theta=1.0, facets=False, **solver_params):
    from waterscapes_amd import (MPETProblem, MPETSolver, Constant, Expression, UnitCubeMesh,
                                 CompiledSubDomain, interpolate, assign)
    A = params["J"]
    u, p, f, g = mms.exact_solutions(params)
    mesh = UnitCubeMesh(n, n, n)
    time = Constant(0.0)
    problem = MPETProblem(mesh, time, params=params)
    problem.f = Expression(tuple(cstr(e) for e in f), t=time, degree=3)
    problem.g = [Expression(cstr(gi), t=time, degree=3) for gi in g]
    problem.u_bar = Expression(tuple(cstr(e) for e in u), t=time, degree=3)
    problem.p_bar = [Expression(cstr(p[i]), t=time, degree=3) for i in range(1, A + 1)]
    p_ex = [Expression(cstr(p[i]), t=time, degree=3) for i in range(1, A + 1)]
    CompiledSubDomain("on_boundary").mark(problem.momentum_boundary_markers, 0)
    for i in range(A):
        CompiledSubDomain("on_boundary").mark(problem.continuity_boundary_markers[i], 0)
    if facets:
        # traction on x = 1 (u), flux on y = 1 (p_1), Robin on z = 1 (p_2)
        CompiledSubDomain("on_boundary && near(x[0], 1.0)").mark(
            problem.momentum_boundary_markers, 1)
        CompiledSubDomain("on_boundary && near(x[1], 1.0)").mark(
            problem.continuity_boundary_markers[0], 1)
        CompiledSubDomain("on_boundary && near(x[2], 1.0)").mark(
            problem.continuity_boundary_markers[1], 2)
        problem.s = Expression(("1.0 + x[1]*t", "x[2]", "-0.5"), t=time, degree=1)
        problem.I = [Expression("0.3*x[0] + t", t=time, degree=1), Constant(0.0)]
        problem.beta = [Constant(0.0), Expression("2.0 + x[0]", t=time, degree=1)]
        problem.p_robin = [Constant(0.0), Expression("0.1*x[1] + 0.2*t", t=time, degree=1)]
    sp_ = dict(dt=dt, theta=theta, T=T)
    sp_.update(solver_params)
    solver = MPETSolver(problem, sp_)
    VP = solver.up_.function_space()
    assign(solver.up_.sub(0), interpolate(problem.u_bar, VP.sub(0).collapse()))
    for i in range(A):
        assign(solver.up_.sub(i + 1), interpolate(p_ex[i], VP.sub(i + 1).collapse()))
    return problem, solver, (u, p, f, g)


def oracle_problem(solver, params, sym, dt, theta, T, facets=False):
    u, p, f, g = sym
    A = params["J"]
    space = solver.up_.function_space()
    m = oracle_mesh(space)
    prob = solver.problem
    mm = np.asarray(prob.momentum_boundary_markers.array(), dtype=np.uint64)
    cm = [np.asarray(c.array(), dtype=np.uint64) for c in prob.continuity_boundary_markers]
    kw = {}
    if facets:
        kw = dict(s=lambda X, t, nrm: np.stack([1.0 + X[:, 1] * t, X[:, 2],
                                                np.full(X.shape[0], -0.5)], 1), s_degree=1,
                  I=[lambda X, t: 0.3 * X[:, 0] + t, None], I_degree=1,
                  beta=[lambda X, t: np.zeros(X.shape[0]), lambda X, t: 2.0 + X[:, 0]],
                  p_robin=[lambda X, t: np.zeros(X.shape[0]),
                           lambda X, t: 0.1 * X[:, 1] + 0.2 * t], robin_degree=1)
    return O.OracleProblem(
        m, params, dict(dt=dt, theta=theta, T=T), f=mms.lambdify_vec(f),
        g=[mms.lambdify(gi) for gi in g], u_bar=mms.lambdify_vec(u),
        p_bar=[mms.lambdify(p[j + 1]) for j in range(A)], f_degree=3, g_degree=3,
        momentum_markers=mm, continuity_markers=cm, formulation="two_field", **kw)


def field_errors(space, got, ref):
    out = []
    for i in range(space.num_sub_spaces()):
        idx = space.field_index(i)
        out.append(np.linalg.norm(got[idx] - ref[idx]) / max(np.linalg.norm(ref[idx]), 1e-300))
    return out


def _csr_to_scipy(csr, n):
    import scipy.sparse as sp
    rp, ci, v = (t.cpu().numpy() for t in csr)
    return sp.csr_matrix((v[: rp[-1]], ci[: rp[-1]], rp), shape=(n, n))


@pytest.mark.parametrize("facets", [False, True])
def test_two_field_matrices_match_oracle(facets):
    problem, solver, sym = build(3, facets=facets)
    eng = solver.engine()
    sysd = eng.system
    ref = oracle_problem(solver, CFG1, sym, 0.05, 1.0, 0.1, facets)
    n = sysd.n
    got_A = _csr_to_scipy(sysd.csr, n)
    assert abs(got_A - ref.A_mat).max() <= 1e-12 * abs(ref.A_mat).max()
    got_P = _csr_to_scipy(sysd.expand_csr(1), n)
    assert abs(got_P - ref.P_mat).max() <= 1e-12 * abs(ref.P_mat).max()
    got_B = _csr_to_scipy(sysd.Bcsr, n)
    Bref = ref.B_matrix()
    assert abs(got_B - Bref).max() <= 1e-12 * abs(Bref).max()


@pytest.mark.parametrize("facets", [False, True])
def test_two_field_solve_matches_oracle_direct(facets):
    problem, solver, sym = build(4, facets=facets)
    ref_prob = oracle_problem(solver, CFG1, sym, 0.05, 1.0, 0.1, facets)
    up0 = solver.up_.vector().get_local()
    ref = [(x.copy(), t) for x, t in ref_prob.solve_direct(up0)]
    space = solver.up_.function_space()
    k = 0
    for (up, t) in solver.solve():
        assert abs(t - ref[k][1]) < 1e-14
        errs = field_errors(space, up.vector().get_local(), ref[k][0])
        assert max(errs) <= 1e-8, (k, errs, solver.solver_monitor)
        k += 1
    assert k == len(ref) == 2
    assert np.array_equal(solver.up_.vector().get_local(), solver.up.vector().get_local())
    assert len(solver.up.split()) == CFG1["J"] + 1


@pytest.mark.parametrize("facets", [False, True])
def test_two_field_mixed_matches_oracle(facets):
    """precision="mixed": FP32 MINRES / preconditioner on the FP32 copy of
    the CSR system inside FP64 refinement to outer_tol 1e-12 — the same
    steps as the oracle's LU to 1e-8 per field."""
    problem, solver, sym = build(4, facets=facets, precision="mixed")
    ref_prob = oracle_problem(solver, CFG1, sym, 0.05, 1.0, 0.1, facets)
    up0 = solver.up_.vector().get_local()
    ref = [(x.copy(), t) for x, t in ref_prob.solve_direct(up0)]
    space = solver.up_.function_space()
    k = 0
    for (up, t) in solver.solve():
        errs = field_errors(space, up.vector().get_local(), ref[k][0])
        st = solver.engine().stats[-1]
        print("two-field mixed", facets, k, errs, st.outer_iterations, st.inner_iterations,
              flush=True)
        assert st.outer_iterations >= 1 and st.converged
        assert max(errs) <= 1e-8, (k, errs)
        k += 1
    assert k == len(ref) == 2


def test_two_field_step_api():
    """step(dt) (:317-379): one step from the current time; up and time move,
    up_ does not."""
    problem, solver, sym = build(3)
    ref_prob = oracle_problem(solver, CFG1, sym, 0.05, 1.0, 0.05)
    up0 = solver.up_.vector().get_local()
    ref, t_ref = next(ref_prob.solve_direct(up0))
    solver.step(0.05)
    assert abs(float(problem.time) - t_ref) < 1e-14
    errs = field_errors(solver.up_.function_space(), solver.up.vector().get_local(), ref)
    assert max(errs) <= 1e-8, errs
    assert np.array_equal(solver.up_.vector().get_local(), up0)
