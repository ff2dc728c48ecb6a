"""Drop-in API on the GPU vs the FP64 oracle (cfg1 manufactured solution).

The problem is set up the way the reference test does it
(test_convergence_totalpressuresolver.py:123-172): sympy -> C strings ->
Expression(degree=3), CompiledSubDomain markers, initial conditions through
assign(up_.sub(i), interpolate(...)), then `for (up, t) in solver.solve()`.
Each step's fields must match the oracle's direct solve (row-replaced LU,
:330-411) to 1e-8 relative L2 per field."""
import re

import numpy as np
import pytest
import sympy

from conftest import CFG1, oracle_mesh
from oracle import mpet_oracle as O
from oracle import mms

pytestmark = pytest.mark.gpu


def cstr(e):
    return re.sub(r"\bx([0-2])\b", r"x[\1]", sympy.printing.ccode(e))


def build_problem(n, params=CFG1, precision="fp64", direct=False, T=0.1, dt=0.05, theta=1.0,
                  mesh=None, **solver_params):
    from waterscapes_amd import (MPETProblem, MPETTotalPressureSolver, Constant, Expression,
                                 UnitCubeMesh, CompiledSubDomain, interpolate, assign)
    A = params["J"]
    u, p, f, g = mms.exact_solutions(params)
    mesh = UnitCubeMesh(n, n, n) if mesh is None else mesh
    time = Constant(0.0)
    problem = MPETProblem(mesh, time, params=params)
    problem.f = Expression(tuple(cstr(e) for e in f), t=time, degree=3)
    problem.g = [Expression(cstr(gi), t=time, degree=3) for gi in g]
    problem.u_bar = Expression(tuple(cstr(e) for e in u), t=time, degree=3)
    p_ex = [Expression(cstr(p[i]), t=time, degree=3) for i in range(A + 1)]
    problem.p_bar = [Expression(cstr(p[i]), t=time, degree=3) for i in range(1, A + 1)]
    on_boundary = CompiledSubDomain("on_boundary")
    on_boundary.mark(problem.momentum_boundary_markers, 0)
    for i in range(A):
        on_boundary.mark(problem.continuity_boundary_markers[i], 0)
    sp_ = dict(dt=dt, theta=theta, T=T, testing=False, direct_solver=direct,
               precision=precision)
    sp_.update(solver_params)
    solver = MPETTotalPressureSolver(problem, sp_)
    VP = solver.up_.function_space()
    V = VP.sub(0).collapse()
    assign(solver.up_.sub(0), interpolate(problem.u_bar, V))
    for i in range(A + 1):
        Q = VP.sub(i + 1).collapse()
        assign(solver.up_.sub(i + 1), interpolate(p_ex[i], Q))
    return problem, solver, (u, p, f, g)


def oracle_steps(solver, params, sym, dt, theta, T):
    u, p, f, g = sym
    A = params["J"]
    space = solver.up_.function_space()
    m = oracle_mesh(space)
    mesh = space.mesh
    mk = np.full(mesh.num_facets(), 2**63 - 1, dtype=np.uint64)
    mk[mesh.exterior] = 0
    prob = O.OracleProblem(
        m, params, dict(dt=dt, theta=theta, T=T), f=mms.lambdify_vec(f),
        g=[mms.lambdify(gi) for gi in g], u_bar=mms.lambdify_vec(u),
        p_bar=[mms.lambdify(p[j + 1]) for j in range(A)], f_degree=3, g_degree=3,
        momentum_markers=mk, continuity_markers=[mk] * A)
    up0 = solver.up_.vector().get_local()
    return [(x.copy(), t) for x, t in prob.solve_direct(up0)]


def field_errors(space, got, ref):
    out = []
    for i in range(space.A + 2):
        idx = space.field_index(i)
        out.append(np.linalg.norm(got[idx] - ref[idx]) / max(np.linalg.norm(ref[idx]), 1e-300))
    return out


# tolerances at which the stopping tests imply <= 1e-8 per-field error
CASES = [("fp64", False, dict(krylov_rtol=1e-12)), ("mixed", False, dict(outer_tol=1e-11)),
         ("fp64", True, {})]


@pytest.mark.parametrize("precision,direct,extra", CASES)
def test_api_mms_matches_oracle(precision, direct, extra):
    params = CFG1
    problem, solver, sym = build_problem(4, params, precision, direct, **extra)
    ref = oracle_steps(solver, params, sym, 0.05, 1.0, 0.1)
    space = solver.up_.function_space()
    k = 0
    for (up, t) in solver.solve():
        assert abs(t - ref[k][1]) < 1e-14
        errs = field_errors(space, up.vector().get_local(), ref[k][0])
        assert max(errs) <= 1e-8, (precision, direct, k, errs, solver.solver_monitor)
        k += 1
    assert k == len(ref) == 2
    assert len(solver.solver_monitor["niter"]) == 2
    # up_ follows up after the yield (reference :497-499)
    assert np.array_equal(solver.up_.vector().get_local(), solver.up.vector().get_local())
    # split() gives (u, p0, p1, p2)
    parts = solver.up.split()
    assert len(parts) == params["J"] + 2


def test_direct_after_iterative_on_one_solver():
    """solve_direct() after solve_iterative() on the same solver runs the
    direct path's FP64 solve (the engine is rebuilt for it, and the initial
    guess history starts afresh): each step matches the oracle's LU to 1e-8
    per field; the iterative run before it used the mixed engine."""
    params = CFG1
    problem, solver, sym = build_problem(4, params, "mixed", False, outer_tol=1e-11)
    up0 = solver.up_.vector().get_local().copy()
    ref = oracle_steps(solver, params, sym, 0.05, 1.0, 0.1)
    it_states = [up.vector().get_local().copy() for up, t in solver.solve_iterative()]
    assert solver.engine(direct=False).krylov.cfg.mode == "mixed"
    # rewind the problem to the initial state and run the direct path
    problem.time.assign(0.0)
    solver.up_.vector()[:] = up0
    space = solver.up_.function_space()
    k = 0
    for up, t in solver.solve_direct():
        errs = field_errors(space, up.vector().get_local(), ref[k][0])
        assert max(errs) <= 1e-8, (k, errs)
        k += 1
    assert k == len(ref) == len(it_states) == 2
    eng = solver.engine(direct=True)
    assert eng.krylov.cfg.mode == "fp64" and eng.krylov.cfg.rtol == 1e-12
