"""Failure paths of the solve, surfaced through the drop-in API, the Krylov
handle and the batch (MPET_ERR_NOT_CONVERGED = 3, MPET_ERR_NAN = 4).

Both conditions are detected on the host from the device MINRES state; every
kernel runs to its normal end (a NaN is data, not an index, and an exhausted
iteration budget ends the loop), so these tests exercise the error reporting
only.  The reference raises from PETSc's KSP on non-convergence
(PETScKrylovSolver, mpettotalpressuresolver.py:448, :492-494)."""
import numpy as np
import pytest
import torch

from conftest import CFG1
from test_gpu_api import build_problem, field_errors
from test_gpu_batch import ISOLATION_RTOL, _batch, _make, _run

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,extra", [("fp64", dict(krylov_rtol=1e-12)),
                                             ("mixed", dict(outer_tol=1e-11))])
def test_maxit_one_reports_not_converged(precision, extra):
    from waterscapes_amd._lib import MpetError
    problem, solver, sym = build_problem(3, CFG1, precision, False, maxit=1, **extra)
    with pytest.raises(MpetError) as ei:
        next(iter(solver.solve()))
    assert ei.value.code == 3
    # the handle reports the same status without raising when asked to
    eng = solver.engine()
    x = solver.up_.data.clone()
    st = eng.krylov.solve(eng.b, x, raise_on_fail=False)
    assert st.converged == 0 and 1 <= st.inner_iterations
    assert torch.isfinite(x).all()


@pytest.mark.parametrize("precision,extra", [("fp64", dict(krylov_rtol=1e-12)),
                                             ("mixed", dict(outer_tol=1e-11))])
def test_nan_in_state_reports_nan(precision, extra):
    from waterscapes_amd._lib import MpetError
    problem, solver, sym = build_problem(3, CFG1, precision, False, **extra)
    space = solver.up_.function_space()
    # a free pressure DOF of network 1 (the interior P1 node of the n = 3 cube)
    idx = space.field_index(2)
    dm = solver.engine().system.dmask.cpu().numpy().astype(bool)
    free = idx[~dm[idx]]
    assert free.size
    solver.up_.data[int(free[0])] = float("nan")
    with pytest.raises(MpetError) as ei:
        next(iter(solver.solve()))
    assert ei.value.code == 4


def test_batch_nan_instance_reports_and_others_match_alone():
    """A NaN in instance 0's state: the batch reports MPET_ERR_NAN, and the
    other instances' first steps equal their solo runs."""
    from waterscapes_amd._lib import MpetError
    items = _make([0, 1, 2])
    batch = _batch(items, "mixed")
    s0 = batch.solvers[0]
    dm = batch.engines[0].system.dmask.cpu().numpy().astype(bool)
    free = np.nonzero(~dm)[0]
    s0.up_.data[int(free[-1])] = float("nan")   # a view into the packed state
    with pytest.raises(MpetError) as ei:
        batch.step()
    assert ei.value.code == 4
    st = ei.value.stats
    assert st[0].converged == 0 and st[1].converged and st[2].converged
    for k in (1, 2):
        alone = _make([k])
        ref = _run(_batch(alone, "mixed"))[0][0][0]
        got = batch.solvers[k].up.vector().get_local()
        space = alone[0][2].up_.function_space()
        errs = field_errors(space, got, ref)
        assert max(errs) <= ISOLATION_RTOL, (k, errs)
    batch.close()


def test_batch_maxit_reports_not_converged():
    from waterscapes_amd._lib import MpetError
    items = _make([0, 1])
    batch = _batch(items, "fp64", maxit=1)
    with pytest.raises(MpetError) as ei:
        batch.step()
    assert ei.value.code == 3
    stats = batch.step(raise_on_fail=False)
    assert all(stats[k].converged == 0 for k in range(2))
    for s in batch.solvers:
        assert torch.isfinite(s.up.data).all()
    batch.close()
