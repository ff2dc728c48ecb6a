"""Autotuned launch shapes on the GPU: the picks are valid shapes, they are
reused from the cache by a second solver of the same shape without timing
anything (also from the cache file in a new cache object), and a tuned
solve gives the untuned solve's fields to the solver tolerance.  Same for a
batch of uneven instances."""
import numpy as np
import pytest

from test_gpu_api import field_errors

pytestmark = pytest.mark.gpu


def _solver(tune, n=16, precision="mixed"):
    import bench
    s = bench.make_solver(n, 4, precision, autotune=tune, T=2 * 0.0125,
                          outer_tol=1e-11)
    return s


def _run(s):
    out = []
    for up, t in s.solve():
        out.append(up.vector().get_local().copy())
    return out


def test_solver_picks_are_valid_reused_and_accurate(tmp_path, monkeypatch):
    from waterscapes_amd import autotune as at
    cache = at.TuneCache(path=str(tmp_path / "tune.json"))
    monkeypatch.setattr(at, "_default_cache", cache)
    tuned = _solver(True)
    tun = tuned.engine().krylov.tuning
    assert 1 in tun and cache.timed > 0
    # the inner FP32 MINRES runs the cell-wise operator (its row kernel's
    # lanes are tuned), and so do the FP64 residuals (same launch shapes):
    # no structured operator is tuned
    assert tun[1]["operator"] is None and "op64" not in tun
    cell = tun["cell"]
    assert cell[0] in at.LANES_CELL_U and cell[1] in at.LANES_CELL_P and cell[2] in at.XCD_CHUNKS
    d64 = tuned.engine().krylov.cell_desc64
    assert d64.fp64 == 1 and [d64.lanes_u, d64.lanes_p, d64.xcd_chunk] == cell
    for rows in (tun[1]["u_levels"], tun[1]["p_levels"]):
        for r in rows:
            assert all(v == 0 or v in at.LPR_ALL for v in r)
    # a second solver of the same shape: every pick from the cache
    timed = cache.timed
    again = _solver(True)
    assert again.engine().krylov.tuning == tun
    assert cache.timed == timed
    # and from the file in a fresh cache object
    monkeypatch.setattr(at, "_default_cache", at.TuneCache(path=str(tmp_path / "tune.json")))
    third = _solver(True)
    assert third.engine().krylov.tuning == tun
    assert at._default_cache.timed == 0
    plain = _solver(False)
    assert plain.engine().krylov.tuning == {}
    space = tuned.up_.function_space()
    for a, b in zip(_run(tuned), _run(plain)):
        assert max(field_errors(space, a, b)) <= 1e-8


def test_batch_tuning_keeps_results(tmp_path, monkeypatch):
    from waterscapes_amd import autotune as at
    from test_gpu_batch import _make, _run as run_batch, _batch
    monkeypatch.setattr(at, "_default_cache", at.TuneCache(path=str(tmp_path / "t.json")))
    items = _make([0, 1, 2])
    b1 = _batch(items, "mixed", autotune=True)
    assert b1.tuning is not None and len(b1.tuning["u_levels"]) >= 1
    r1 = run_batch(b1)
    b1.close()
    items2 = _make([0, 1, 2])
    r2 = run_batch(_batch(items2, "mixed"))
    for k, it in enumerate(items):
        space = it[2].up_.function_space()
        for (a, _), (b, _) in zip(r1[k], r2[k]):
            assert max(field_errors(space, a, b)) <= 1e-8
