"""Batched stepping of uneven instances on the GPU.

Three manufactured-solution problems with different meshes, network counts,
material parameters and end times are packed into one batch.  Checked:
  * the packed operator and preconditioner act on every instance's block
    like that instance's own operator (FP64 oracle matrix) and the numpy
    restatement of its V-cycles;
  * every instance's fields at every step match the FP64 oracle's direct
    solve (1e-8 relative L2 per field), for FP64 and mixed batches;
  * isolation: an instance's fields are the same alone and in the mix
    (ISOLATION_RTOL relative L2 per field);
  * instances marked inactive are left untouched.
"""
import numpy as np
import pytest

from conftest import CFG1, oracle_mesh
from oracle import mpet_oracle as O
from test_gpu_api import build_problem, field_errors, oracle_steps
from test_gpu_solver import _ref_precond

pytestmark = pytest.mark.gpu

P_A1 = dict(J=1, E=5e3, nu=0.3, alpha=(0.4,), c=(1e-3,), K=(1e-2,), S=((0.0,),))
P_A3 = dict(J=3, E=2e4, nu=0.45, alpha=(0.3, 0.4, 0.5), c=(1e-4, 2e-4, 1e-3),
            K=(1e-3, 1e-4, 1e-2), S=((0.0, 1e-2, 0.0), (1e-2, 0.0, 3e-3), (0.0, 3e-3, 0.0)))
# (mesh n, params, T): dt = 0.05 for all, so instance 2 stops after one step
INSTANCES = [(3, CFG1, 0.1), (4, P_A1, 0.1), (2, P_A3, 0.05)]
ACC_RTOL = 1e-8          # per-field relative L2 error vs the oracle
ISOLATION_RTOL = 1e-9    # per-field relative L2, alone vs in the mix


def _make(which):
    out = []
    for i in which:
        n, params, T = INSTANCES[i]
        prob, solver, sym = build_problem(n, params, "fp64", False, T=T, dt=0.05)
        out.append((params, prob, solver, sym, T))
    return out


def _batch(items, precision, **kw):
    from waterscapes_amd.batch import MPETBatch
    defaults = dict(rtol=1e-12) if precision == "fp64" else dict(outer_tol=1e-11)
    defaults.update(kw)
    return MPETBatch([it[2] for it in items], precision=precision, **defaults)


def _run(batch):
    """All steps of the batch: per instance the list of (fields, t)."""
    res = [[] for _ in batch.solvers]
    for stepped, stats in batch.solve():
        for k in stepped:
            assert stats[k].converged
            res[k].append((batch.solvers[k].up.vector().get_local().copy(), batch.t[k]))
    return res


def test_batch_operator_and_preconditioner_blocks():
    import torch
    items = _make([0, 1, 2])
    batch = _batch(items, "fp64")
    rng = np.random.default_rng(4)
    x_i = torch.tensor(rng.standard_normal(batch.n), device="cuda")
    y = batch.unpack(batch.apply(batch.pack(x_i))).cpu().numpy()
    z = batch.unpack(batch.precond(batch.pack(x_i), dtype=0)).cpu().numpy()
    xh = x_i.cpu().numpy()
    for k, (params, prob, solver, sym, T) in enumerate(items):
        a, b = int(batch.n_off[k]), int(batch.n_off[k + 1])
        eng = batch.engines[k]
        m = oracle_mesh(eng.space)
        A = O.assemble_matrix(m["coords"], m["cells"], m["cell_p2"], m["cell_p1"], m["nP2"],
                              m["nP1"], params, 0.05, 1.0, which="A")
        ref = A @ xh[a:b]
        ref[eng.system.dmask.cpu().numpy().astype(bool)] = 0.0
        assert np.abs(y[a:b] - ref).max() <= 1e-12 * np.abs(ref).max(), k
        r = xh[a:b].copy()
        r[eng.system.dmask.cpu().numpy().astype(bool)] = 0.0
        zr = _ref_precond(eng.krylov, r)
        zg = z[a:b]
        zg[eng.system.dmask.cpu().numpy().astype(bool)] = 0.0
        assert np.abs(zg - zr).max() <= 1e-9 * np.abs(zr).max(), k
    batch.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_matches_oracle(precision):
    items = _make([0, 1, 2])
    refs = [oracle_steps(it[2], it[0], it[3], 0.05, 1.0, it[4]) for it in items]
    batch = _batch(items, precision)
    res = _run(batch)
    for k, it in enumerate(items):
        space = it[2].up_.function_space()
        assert len(res[k]) == len(refs[k]) == round(it[4] / 0.05), k
        for (got, t), (ref, tr) in zip(res[k], refs[k]):
            assert abs(t - tr) < 1e-14
            errs = field_errors(space, got, ref)
            assert max(errs) <= ACC_RTOL, (precision, k, errs)
        mon = it[2].solver_monitor
        assert len(mon["niter"]) == len(refs[k]) and min(mon["niter"]) > 0
    batch.close()


def test_batch_isolation():
    """Each instance gives the same fields alone and in the mix."""
    mixed = _make([0, 1, 2])
    res_mix = _run(_batch(mixed, "mixed"))
    for k in range(3):
        alone = _make([k])
        res_alone = _run(_batch(alone, "mixed"))[0]
        space = alone[0][2].up_.function_space()
        assert len(res_alone) == len(res_mix[k])
        for (a, ta), (m, tm) in zip(res_alone, res_mix[k]):
            assert ta == tm
            errs = field_errors(space, m, a)
            assert max(errs) <= ISOLATION_RTOL, (k, errs)


def test_batch_inactive_instances_untouched():
    import torch
    items = _make([0, 1, 2])
    batch = _batch(items, "mixed")
    before = [s.up.data.clone() for s in batch.solvers]
    t0 = list(batch.t)
    stats = batch.step(active=[1, 0, 1])
    assert stats[1].inner_iterations == 0 and stats[1].converged == 0
    assert stats[0].converged and stats[2].converged
    assert torch.equal(batch.solvers[1].up.data, before[1])
    assert batch.t[1] == t0[1] and batch.t[0] == pytest.approx(t0[0] + 0.05)
    assert not torch.equal(batch.solvers[0].up.data, before[0])
    batch.close()
