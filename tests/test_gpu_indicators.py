"""A-posteriori error indicators (mpet/adaptivempetsolver.py:82-222) on the
GPU vs the oracle: random two-field states and polynomial / trigonometric
sources on a box, every cell's zeta_p, zeta_u, zeta_u_dt and R4p part to
1e-11 relative, and the reference's estimator bookkeeping and markers on
the result."""
import numpy as np
import pytest

from conftest import oracle_mesh
from oracle import mpet_oracle as O

pytestmark = pytest.mark.gpu

PARAMS = dict(J=2, E=800.0, nu=0.3, alpha=(0.7, 0.4), c=(1e-2, 2e-2), K=(0.3, 0.05),
              S=((0.0, 0.15), (0.15, 0.0)))


def test_indicators_match_oracle():
    import torch
    from waterscapes_amd import MPETProblem, MPETSolver, Constant, Expression, BoxMesh
    from waterscapes_amd.adaptive import ErrorEstimator, dorfler_mark, maximal_mark
    mesh = BoxMesh((0.0, 0.0, 0.0), (1.0, 1.5, 0.7), 3, 2, 2)
    time = Constant(0.4)
    problem = MPETProblem(mesh, time, params=dict(PARAMS))
    problem.f = Expression(("x[0]*x[1] + t", "sin(x[2])", "x[0]*x[0]"), t=time, degree=2)
    problem.g = [Expression("x[1] - t", t=time, degree=1), Expression("x[0]*x[2]", degree=2)]
    f_ = Expression(("x[0]*x[1]", "cos(x[2])", "x[1]"), degree=2)
    solver = MPETSolver(problem, dict(dt=0.05, theta=1.0, T=0.1))
    n = solver.up.function_space().dim()
    rng = np.random.default_rng(5)
    x = rng.standard_normal(n)
    xp = rng.standard_normal(n)
    solver.up.data.copy_(torch.as_tensor(x))
    solver.up_.data.copy_(torch.as_tensor(xp))
    est = ErrorEstimator(problem, solver, f_)
    print("indicators: cells", mesh.num_cells(), "interior facets", est.nF, flush=True)
    got = est.indicators()
    m = oracle_mesh(solver.up.function_space())
    ref = O.error_indicators(
        m, PARAMS, x, xp, 0.05,
        f=lambda X, t: np.stack([X[:, 0] * X[:, 1] + t, np.sin(X[:, 2]), X[:, 0] ** 2], 1),
        fp=lambda X, t: np.stack([X[:, 0] * X[:, 1], np.cos(X[:, 2]), X[:, 1]], 1),
        g=lambda X, t: np.stack([X[:, 1] - t, X[:, 0] * X[:, 2]], 1),
        f_degree=2, g_degree=2, t=0.4)
    for name, a, b in zip(("zeta_p", "zeta_u", "zeta_u_dt", "r4"), got, ref):
        err = np.abs(a - b).max() / np.abs(b).max()
        print(name, err, flush=True)
        assert err < 1e-11
    eta_Ks = est.estimate_error_at_t()
    est.add_to_indicators(eta_Ks)
    ind = est.compute_error_indicators()
    want = ref[1] + 0.05 * np.sqrt(ref[2]) + np.sqrt(0.05 * ref[0])
    assert np.allclose(ind, want, rtol=1e-11, atol=0)
    eta, etas = est.compute_error_estimate()
    assert abs(etas[0] - np.sqrt(0.05 * ref[0].sum())) < 1e-11 * etas[0]
    assert abs(etas[3] - np.sqrt(0.05 * ref[3].sum())) < 1e-11 * etas[3]
    assert maximal_mark(ind, 0.25).sum() == int(np.ceil(0.25 * ind.size))
    mk = dorfler_mark(ind, 0.5)
    assert ind[mk].sum() >= 0.5 * ind.sum() and ind[mk].min() >= ind[~mk].max()
