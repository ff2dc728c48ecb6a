"""Autotuner logic on the CPU: candidate order, the incumbent's margin, the
cache's reuse across instances and processes (a fake timer stands in for
the device)."""
import json

from waterscapes_amd import autotune as at


def test_pick_keeps_incumbent_within_margin():
    cache = at.TuneCache(path="/nonexistent-dir/x.json")
    times = {1: 1.00, 2: 0.99, 4: 1.20}
    best, _ = at._pick((1, 2, 4), 1, times.get, 0.02, cache)
    assert best == 1                       # 1% faster is within the margin
    times[2] = 0.95
    best, _ = at._pick((1, 2, 4), 1, times.get, 0.02, cache)
    assert best == 2
    assert cache.timed == 6


def test_lane_candidates():
    assert at._lpr_cands(3, 8) == (1, 4, 8, 16, 32, 64)
    assert at._lpr_cands(1, 4) == (1, 2, 4, 8, 16, 32)
    # three columns: one thread per row or at least one lane per component
    assert all(c == 1 or c >= 4 for c in at._lpr_cands(3, 4))
    assert at._lpr_cands(3, 64) == (8, 16, 32, 64)


def test_cache_round_trip(tmp_path):
    p = str(tmp_path / "tune.json")
    c1 = at.TuneCache(path=p)
    key = at.level_key(1, 3, "A", 912673, 912673, 23518878, True)
    c1.put(key, 8)
    c2 = at.TuneCache(path=p)
    assert c2.get(key) == 8
    d = json.load(open(p))
    assert d["arch"] == at.ARCH and d["picks"][key] == 8
    # another architecture's file is ignored
    json.dump({"arch": "gfx942", "picks": {key: 4}}, open(p, "w"))
    assert at.TuneCache(path=p).get(key) is None
