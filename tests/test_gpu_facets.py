"""Exterior-facet terms on the GPU against the FP64 oracle: traction s on
ds_m(1), network flux I_1 on ds_1(1) and a Robin condition (beta_2,
p_robin_2) on ds_2(2), all on the face x = 1 (Dirichlet elsewhere).

  * the assembled A (Robin term folded into network 2's stiffness) equals
    the oracle's entrywise, and P keeps no Robin term;
  * the right-hand side (theta = 0.5, so the (1 - theta) Robin part is
    active) equals the oracle's;
  * whole steps through the drop-in API match the oracle's direct solve, and
    the steady linear-pressure known answer is reproduced."""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import oracle_mesh
from oracle import mpet_oracle as O
from oracle import mms
from test_gpu_api import cstr, field_errors
from test_oracle import UNIT, _facet_mms_problem

pytestmark = pytest.mark.gpu


def _product_problem(n, params, theta, dt, T, sym, pexprs):
    """The same problem as _facet_mms_problem through the product API."""
    from waterscapes_amd import (MPETProblem, MPETTotalPressureSolver, Constant, Expression,
                                 UnitCubeMesh, CompiledSubDomain, FacetNormal)
    u, p, f, g = sym
    A = int(params["J"])
    sigma, grads = mms.boundary_data(params, u, p, 2.0)
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
    K = params["K"]
    problem.I = [Expression(cstr(-K[0] * grads[0][0]), t=time, degree=3), Constant(0.0)]
    problem.beta = [Constant(0.0), Constant(2.0)]
    problem.p_robin = [Constant(0.0),
                       Expression(cstr(p[2] + K[1] * grads[1][0] / 2.0), t=time, degree=3)]
    onb = CompiledSubDomain("on_boundary")
    right = CompiledSubDomain("on_boundary && near(x[0], 1.0)")
    onb.mark(problem.momentum_boundary_markers, 0)
    right.mark(problem.momentum_boundary_markers, 1)
    for i in range(A):
        onb.mark(problem.continuity_boundary_markers[i], 0)
    right.mark(problem.continuity_boundary_markers[0], 1)
    right.mark(problem.continuity_boundary_markers[1], 2)
    return problem


def _solver(problem, theta, dt, T, **kw):
    from waterscapes_amd import MPETTotalPressureSolver
    sp_ = dict(dt=dt, theta=theta, T=T, direct_solver=False, precision="fp64",
               krylov_rtol=1e-12)
    sp_.update(kw)
    return MPETTotalPressureSolver(problem, sp_)


def _csr_host(t):
    rp, ci, vals = t
    n = rp.numel() - 1
    return sp.csr_matrix((vals.cpu().numpy(), ci.cpu().numpy().astype(np.int64),
                          rp.cpu().numpy()), shape=(n, n))


def _setup(n=3, theta=0.5, dt=0.3, T=0.6):
    mesh, space, m, prob, sym = _facet_mms_problem(n, UNIT, theta=theta, dt=dt, T=T)
    problem = _product_problem(n, UNIT, theta, dt, T, sym, None)
    return m, prob, sym, problem


def test_robin_matrix_and_rhs_match_oracle():
    import torch
    m, prob, sym, problem = _setup()
    solver = _solver(problem, 0.5, 0.3, 0.6)
    eng = solver.engine()
    s = eng.system
    # A with the Robin term, P without it
    A = _csr_host(s.expand_csr(0))
    P = _csr_host(s.expand_csr(1))
    for G, R in ((A, prob.A_mat), (P, prob.P_mat)):
        R = R.tocsr()
        R.sort_indices()
        assert np.array_equal(G.indptr, R.indptr) and np.array_equal(G.indices, R.indices)
        scale = np.abs(R.data).max()
        assert np.abs(G.data - R.data).max() <= 1e-12 * scale
    # right-hand side at an arbitrary state and time
    rng = np.random.default_rng(7)
    up_prev = rng.standard_normal(s.n)
    t_n = 0.3
    b_ref, t_theta, t_np1 = prob.rhs(up_prev.copy(), t_n)
    t_got = eng.assemble_rhs(torch.tensor(up_prev, device="cuda"), t_n)
    b = eng.b.cpu().numpy()
    assert abs(t_got - t_np1) < 1e-15
    assert np.abs(b - b_ref).max() <= 1e-12 * np.abs(b_ref).max()


@pytest.mark.parametrize("precision,extra", [("fp64", {}), ("mixed", dict(outer_tol=1e-11))])
def test_facet_steps_match_oracle(precision, extra):
    import torch
    m, prob, sym, problem = _setup(theta=1.0, dt=0.5, T=1.0)
    u, p, f, g = sym
    solver = _solver(problem, 1.0, 0.5, 1.0, precision=precision, **extra)
    lay = prob.lay
    up0 = np.zeros(lay.n)
    up0[: lay.nu] = mms.lambdify_vec(u)(m["x2"], 0.0).reshape(-1)
    for j in range(3):
        up0[lay.pdofs(np.arange(m["nP1"]), j)] = mms.lambdify(p[j])(m["coords"], 0.0)
    ref = [(x.copy(), t) for x, t in prob.solve_direct(up0.copy())]
    solver.up_.data.copy_(torch.tensor(up0, device="cuda"))
    space = solver.up_.function_space()
    k = 0
    for up, t in solver.solve():
        assert abs(t - ref[k][1]) < 1e-14
        errs = field_errors(space, up.vector().get_local(), ref[k][0])
        assert max(errs) <= 1e-8, (precision, k, errs)
        k += 1
    assert k == len(ref) == 2
