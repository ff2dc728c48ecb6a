"""theta != 1 through the drop-in API on the GPU, with exchange S != 0.

The theta-weighted stiffness and exchange entries of A (dt*theta) and of the
up_-linear part of b (dt*(1 - theta)), and the t_theta time level of the
sources and fluxes (mpettotalpressuresolver.py:239-241, :270-276, :467-476),
checked step by step against the FP64 oracle's direct solve, whose theta
terms tests/test_oracle_theta.py pins:

  * the three-network known answer (pressures linear in space and time,
    flux face x = 1), theta = 0.5 and 0.3: every step equals the oracle and
    the exact nodal values;
  * the reference's Crank-Nicolson experiment in 3D (S = ((1,1),(1,1)),
    traction face x = 1, test_convergence_totalpressuresolver.py:103-172);
  * the cfg1 material set (S_12 = 1e-2) at theta = 0.5.
fp64, mixed and direct paths, <= 1e-8 relative L2 per field and step."""
import numpy as np
import pytest
import torch

from conftest import CFG1
from oracle import mms
from test_gpu_api import build_problem, cstr, field_errors, oracle_steps  # noqa: F401
from test_oracle_theta import (KA_PARAMS, UNIT, _exact_at, known_answer_fields,
                               known_answer_problem, reference_theta_fields,
                               reference_theta_problem)

pytestmark = pytest.mark.gpu

CASES = [("fp64", False, dict(krylov_rtol=1e-12)), ("mixed", False, dict(outer_tol=1e-11)),
         ("fp64", True, {})]


def _solver(problem, theta, dt, T, precision, direct, extra):
    from waterscapes_amd import MPETTotalPressureSolver
    sp_ = dict(dt=dt, theta=theta, T=T, direct_solver=direct, precision=precision)
    sp_.update(extra)
    return MPETTotalPressureSolver(problem, sp_)


def _known_answer_product(n, params=KA_PARAMS):
    from waterscapes_amd import (MPETProblem, Constant, Expression, UnitCubeMesh,
                                 CompiledSubDomain)
    u, p, f, g, flux = known_answer_fields(params)
    A = int(params["J"])
    mesh = UnitCubeMesh(n, n, n)
    time = Constant(0.0)
    problem = MPETProblem(mesh, time, params=params)
    problem.f = Expression(tuple(cstr(e) for e in f), t=time, degree=1)
    problem.g = [Expression(cstr(gi), t=time, degree=1) for gi in g]
    problem.p_bar = [Expression(cstr(p[i]), t=time, degree=1) for i in range(1, A + 1)]
    problem.I = [Expression(cstr(q), t=time, degree=1) for q in flux]
    onb = CompiledSubDomain("on_boundary")
    right = CompiledSubDomain("on_boundary && near(x[0], 1.0)")
    onb.mark(problem.momentum_boundary_markers, 0)
    for i in range(A):
        onb.mark(problem.continuity_boundary_markers[i], 0)
        right.mark(problem.continuity_boundary_markers[i], 1)
    return problem


def _field_errors(space, got, ref):
    """Relative L2 error per field; a field whose reference vanishes (u of
    the known answer) is measured relative to the whole state."""
    whole = np.linalg.norm(ref)
    out = []
    for i in range(space.A + 2):
        idx = space.field_index(i)
        out.append(np.linalg.norm(got[idx] - ref[idx]) / max(np.linalg.norm(ref[idx]), whole))
    return out


def _run_against(solver, ref, up0, exact=None):
    solver.up_.data.copy_(torch.tensor(up0, device="cuda"))
    space = solver.up_.function_space()
    k = 0
    for up, t in solver.solve():
        assert abs(t - ref[k][1]) < 1e-14
        got = up.vector().get_local()
        errs = _field_errors(space, got, ref[k][0])
        assert max(errs) <= 1e-8, (k, errs, solver.solver_monitor)
        if exact is not None:
            ex = exact(t)
            assert np.abs(got - ex).max() <= 1e-8 * np.abs(ex).max(), k
        k += 1
    assert k == len(ref)


@pytest.mark.parametrize("theta", [0.5, 0.3])
@pytest.mark.parametrize("precision,direct,extra", CASES)
def test_theta_known_answer_matches_oracle(theta, precision, direct, extra):
    dt, T = 0.25, 0.75
    mesh, space, m, prob, sym = known_answer_problem(2, theta, dt, T)
    up0 = _exact_at(prob, m, sym, 0.0)
    ref = [(x.copy(), t) for x, t in prob.solve_direct(up0.copy())]
    problem = _known_answer_product(2)
    solver = _solver(problem, theta, dt, T, precision, direct, extra)
    _run_against(solver, ref, up0, exact=lambda t: _exact_at(prob, m, sym, t))


def _reference_theta_product(n, params=UNIT):
    from waterscapes_amd import (MPETProblem, Constant, Expression, UnitCubeMesh,
                                 CompiledSubDomain, FacetNormal)
    u, p, f, g = reference_theta_fields(params)
    sigma, _ = mms.boundary_data(params, u, p, 1.0)
    A = int(params["J"])
    mesh = UnitCubeMesh(n, n, n)
    time = Constant(0.0)
    problem = MPETProblem(mesh, time, params=params)
    problem.f = Expression(tuple(cstr(e) for e in f), t=time, degree=3)
    problem.g = [Expression(cstr(gi), t=time, degree=3) for gi in g]
    problem.u_bar = Expression(tuple(cstr(e) for e in u), t=time, degree=3)
    problem.p_bar = [Expression(cstr(p[i]), t=time, degree=3) for i in range(1, A + 1)]
    sig = Expression(tuple(tuple(cstr(sigma[i][j]) for j in range(3)) for i in range(3)),
                     t=time, degree=3)
    problem.s = sig * FacetNormal(mesh)
    onb = CompiledSubDomain("on_boundary")
    right = CompiledSubDomain("on_boundary && near(x[0], 1.0)")
    onb.mark(problem.momentum_boundary_markers, 0)
    right.mark(problem.momentum_boundary_markers, 1)
    for i in range(A):
        onb.mark(problem.continuity_boundary_markers[i], 0)
    return problem


@pytest.mark.parametrize("precision,direct,extra", CASES)
def test_reference_theta_half_experiment_matches_oracle(precision, direct, extra):
    """n = 4, M = 2 steps to T = 1 at theta = 0.5."""
    n, M = 4, 2
    mesh, space, m, prob, sym = reference_theta_problem(n, M)
    from test_oracle import _initial
    up0 = _initial(prob, m, sym[0], sym[1])
    ref = [(x.copy(), t) for x, t in prob.solve_direct(up0.copy())]
    problem = _reference_theta_product(n)
    solver = _solver(problem, 0.5, 1.0 / M, 1.0, precision, direct, extra)
    _run_against(solver, ref, up0)


@pytest.mark.parametrize("precision,direct,extra", CASES)
def test_cfg1_material_theta_half_matches_oracle(precision, direct, extra):
    problem, solver, sym = build_problem(4, CFG1, precision, direct, theta=0.5, **extra)
    ref = oracle_steps(solver, CFG1, sym, 0.05, 0.5, 0.1)
    space = solver.up_.function_space()
    k = 0
    for up, t in solver.solve():
        assert abs(t - ref[k][1]) < 1e-14
        errs = field_errors(space, up.vector().get_local(), ref[k][0])
        assert max(errs) <= 1e-8, (precision, direct, k, errs)
        k += 1
    assert k == len(ref) == 2
