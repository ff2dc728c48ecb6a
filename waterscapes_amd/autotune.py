"""Autotuner of the launch shapes of the hot kernels.

The reference leaves launch shapes to hypre and PETSc
(mpet/mpettotalpressuresolver.py:448); here they are the lanes that share
one matrix row:

  * the structured operator (fused MINRES step and mpet_op_apply): lanes
    per P2 node row (u rows) and per P1 node row (p rows) in {8, 16, 32},
    and the run length of row tiles dealt to one XCD (xcd_chunk, 0 or 4);
  * the cell-wise FP32 operator (mpet_cell_op_apply): lanes of its row
    kernel per P2 node row in {4, 8, 16} and per P1 node row in {8, 16},
    and its XCD run length (0 or 4);
  * every AMG level of both hierarchies: lanes per row of the level
    operator (timed as the Chebyshev step), of the prolongator and of the
    restriction (timed as plain products), powers of two up to 64.

Each candidate is timed on the device with HIP events on the current
stream (the mean of `reps` launches, the best of `repeats` such means), in
a fixed order.  A candidate replaces the incumbent (the package's fixed
heuristic) only when it is faster by more than `margin`, so timing noise
leaves the heuristic in place rather than flipping between near-equal
shapes.  Picks are cached by the statistics that determine the kernel's
work (rows, nonzeros, columns, precision, compact storage; the operator's
node-pattern sizes and field count) and the architecture, in memory and in
a JSON file (MPET_TUNE_CACHE, default ~/.cache/waterscapes_amd/), so a
second solver of the same shape reuses them without timing anything.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import threading

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr

ARCH = "gfx950"
LANES_OP = (8, 16, 32)
LANES_CELL_U = (1, 4, 8, 16)
LANES_CELL_P = (8, 16)
XCD_CHUNKS = (0, 4)
LPR_ALL = (1, 2, 4, 8, 16, 32, 64)


class Tuning(C.Structure):
    _fields_ = [("lanes_u", C.c_int32), ("lanes_p", C.c_int32), ("xcd_chunk", C.c_int32),
                ("u_lpr", (C.c_int32 * 3) * 16), ("p_lpr", (C.c_int32 * 3) * 16)]


def _bind():
    L = _lib.lib()
    P = C.c_void_p
    L.mpet_solver_set_tuning.argtypes = [P, C.c_int, P]
    L.mpet_solver_set_tuning.restype = C.c_int
    L.mpet_batch_set_tuning.argtypes = [P, C.c_int, P]
    L.mpet_batch_set_tuning.restype = C.c_int
    L.mpet_level_spmm.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, P, P, P, P, P, P, P, P,
                                  C.c_double, C.c_double, P, C.c_int, P, P, C.c_double, P]
    L.mpet_level_spmm.restype = C.c_int
    return L


class TuneCache:
    """Picks by key, in memory and in a JSON file (one per architecture)."""

    _lock = threading.Lock()

    def __init__(self, path=None):
        self.path = path or os.environ.get("MPET_TUNE_CACHE") or os.path.join(
            os.path.expanduser("~"), ".cache", "waterscapes_amd", f"tune-{ARCH}.json")
        self.data = {}
        self.timed = 0            # candidates timed through this cache (tests)
        try:
            with open(self.path) as f:
                d = json.load(f)
            if d.get("arch") == ARCH:
                self.data = d.get("picks", {})
        except (OSError, ValueError):
            pass

    def get(self, key):
        return self.data.get(key)

    def put(self, key, value):
        with self._lock:
            self.data[key] = value
            try:
                os.makedirs(os.path.dirname(self.path), exist_ok=True)
                tmp = f"{self.path}.{os.getpid()}.tmp"
                with open(tmp, "w") as f:
                    json.dump({"arch": ARCH, "picks": self.data}, f, indent=1, sort_keys=True)
                os.replace(tmp, self.path)
            except OSError:
                pass              # an unwritable cache directory keeps picks in memory


_default_cache = None


def default_cache():
    global _default_cache
    if _default_cache is None:
        _default_cache = TuneCache()
    return _default_cache


def time_launches(fn, reps=10, repeats=3):
    """Mean duration (ms) of fn() over `reps` launches, best of `repeats`.
    The launches are captured once in a graph and the graph is replayed (as
    the solver replays its iterations), so that host launch overhead does
    not swamp the short kernels of the coarse levels; fn must launch on
    torch's current stream, looked up at each call."""
    fn()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(reps):
            fn()
    g.replay()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    best = float("inf")
    for _ in range(repeats):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        g.replay()
        e1.record(st)
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1) / reps)
    del g
    return best


def _pick(cands, default, timer, margin, cache):
    """Candidate with the smallest time; the default wins unless another is
    faster by more than `margin` (relative)."""
    times = {}
    for c in cands:
        times[c] = timer(c)
        if cache is not None:
            cache.timed += 1
    best = min(cands, key=lambda c: times[c])
    if default in times and times[best] >= (1.0 - margin) * times[default]:
        best = default
    return best, times


# ---------------------------------------------------------------------------
# structured operator
# ---------------------------------------------------------------------------
def op_key(sys_, dtype):
    nz = sys_.nnz
    return (f"op|{ARCH}|{dtype}|F{sys_.F}|{sys_.nP2}|{sys_.nP1}|{nz['22']}|{nz['21']}|"
            f"{nz['12']}|{nz['11']}")


def tune_operator(sys_, dtype, cache=None, reps=10, repeats=3, margin=0.02):
    """(lanes_u, lanes_p, xcd_chunk) of the structured operator."""
    cache = cache if cache is not None else default_cache()
    key = op_key(sys_, dtype)
    hit = cache.get(key)
    if hit is not None:
        return tuple(hit)
    L = _lib.lib()
    tdt = torch.float64 if dtype == 0 else torch.float32
    g = torch.Generator(device=sys_.device).manual_seed(0)
    x = torch.randn(sys_.n, generator=g, device=sys_.device, dtype=torch.float64).to(tdt)
    y = torch.empty_like(x)
    default = (16, 16, sys_.xcd_chunk[dtype])

    def timer(c):
        d = sys_.op_desc(dtype, lanes=(c[0], c[1]), xcd_chunk=c[2])
        return time_launches(lambda: check(L.mpet_op_apply(C.byref(d), ptr(x), None, ptr(y),
                                                           _lib.stream_ptr())), reps, repeats)
    cands = [(u, p, xc) for u in LANES_OP for p in LANES_OP for xc in XCD_CHUNKS]
    best, _ = _pick(cands, default, timer, margin, cache)
    cache.put(key, list(best))
    return best


def cell_key(sys_):
    nz = sys_.nnz
    return (f"cell|{ARCH}|F{sys_.F}|{sys_.cell_op()['nC']}|{sys_.nP2}|{sys_.nP1}|"
            f"{nz['11']}")


def tune_cell_operator(sys_, cache=None, reps=10, repeats=3, margin=0.02):
    """(lanes_u, lanes_p, xcd_chunk) of the cell-wise operator's row kernel."""
    cache = cache if cache is not None else default_cache()
    key = cell_key(sys_)
    hit = cache.get(key)
    if hit is not None:
        return tuple(hit)
    L = _lib.lib()
    g = torch.Generator(device=sys_.device).manual_seed(0)
    x = torch.randn(sys_.n, generator=g, device=sys_.device, dtype=torch.float64).float()
    y = torch.empty_like(x)
    work = torch.empty(sys_.cell_work_floats(), dtype=torch.float32, device=sys_.device)
    default = tuple(sys_.cell_lanes) + (sys_.cell_xcd_chunk,)

    def timer(c):
        d = sys_.cell_op_desc(lanes=(c[0], c[1]), xcd_chunk=c[2])
        return time_launches(lambda: check(L.mpet_cell_op_apply(C.byref(d), ptr(x), None, ptr(y),
                                                                ptr(work), _lib.stream_ptr())),
                             reps, repeats)
    cands = [(u, p, xc) for u in LANES_CELL_U for p in LANES_CELL_P for xc in XCD_CHUNKS]
    best, _ = _pick(cands, default, timer, margin, cache)
    cache.put(key, list(best))
    return best


# ---------------------------------------------------------------------------
# AMG levels
# ---------------------------------------------------------------------------
def _lpr_cands(nc, heuristic):
    # three-column SpMMs: one thread per row (1) or at least one lane per
    # component (4 and up)
    ok = (lambda v: v == 1 or v >= 4) if nc == 3 else (lambda v: True)
    return tuple(v for v in LPR_ALL if ok(v) and heuristic / 8 <= v <= heuristic * 8)


def level_key(dtype, nc, which, rows, cols, nnz, compact):
    return f"lvl|{ARCH}|{dtype}|nc{nc}|{which}|{rows}|{cols}|{nnz}|{int(bool(compact))}"


def tune_hierarchy(levels_host, levels_dev, nc, dtype, heuristic, mask0=None, cache=None,
                   reps=10, repeats=3, margin=0.02, device="cuda"):
    """Lanes per row [[A, P, R] per level] of one hierarchy.  heuristic(l,
    which) gives the fixed pick used as the incumbent."""
    cache = cache if cache is not None else default_cache()
    L = _bind()
    tdt = torch.float64 if dtype == 0 else torch.float32
    out = []
    for l, (lev, ld) in enumerate(zip(levels_host, levels_dev)):
        n = int(lev["n"])
        ncols = int(lev["nc"])
        row = [0, 0, 0]
        mats = [("A", ld["A_rp"], ld["A_ci"], ld["A_v"], n, n)]
        if ncols:
            mats += [("P", ld["P_rp"], ld["P_ci"], ld["P_v"], n, ncols),
                     ("R", ld["R_rp"], ld["R_ci"], ld["R_v"], ncols, n)]
        for j, (which, rp, ci, v, rows, cols) in enumerate(mats):
            if l == len(levels_host) - 1 and which == "A":
                continue          # the coarsest level is a dense solve
            compact = which == "A" and "A_v16" in ld
            nnz = int(ci.numel())
            key = level_key(dtype, nc, which, rows, cols, nnz, compact)
            hit = cache.get(key)
            if hit is not None:
                row[j] = int(hit)
                continue
            x = torch.ones(nc * cols, device=device, dtype=tdt)
            y = torch.empty(nc * rows, device=device, dtype=tdt)
            mode = 3 if which == "A" else 0
            b = torch.ones(nc * rows, device=device, dtype=tdt) if mode == 3 else None
            d = torch.zeros(nc * rows, device=device, dtype=tdt) if mode == 3 else None
            dinv = ld["dinv"] if mode == 3 else None
            m0 = mask0 if (l == 0 and which != "R") else None
            v16 = ld.get("A_v16") if compact else None
            dci = ld.get("A_dci") if compact else None
            vs = float(ld.get("vscale", 1.0)) if compact else 1.0

            def timer(lpr, rp=rp, ci=ci, v=v, x=x, b=b, y=y, dinv=dinv, d=d, m0=m0, v16=v16,
                      dci=dci, vs=vs, rows=rows, mode=mode):
                return time_launches(lambda: check(L.mpet_level_spmm(
                    dtype, nc, mode, rows, ptr(rp), ptr(ci), None if v16 is not None else ptr(v),
                    ptr(x), ptr(b) if b is not None else None, ptr(y),
                    ptr(dinv) if dinv is not None else None, ptr(d) if d is not None else None,
                    0.5, 0.5, ptr(m0) if m0 is not None else None, lpr,
                    ptr(v16) if v16 is not None else None, ptr(dci) if dci is not None else None,
                    vs, _lib.stream_ptr())), reps, repeats)
            dflt = int(heuristic(l, which))
            cands = _lpr_cands(nc, dflt)
            if dflt not in cands:
                cands = cands + (dflt,)
            best, _ = _pick(cands, dflt, timer, margin, cache)
            cache.put(key, int(best))
            row[j] = int(best)
        out.append(row)
    return out


def apply_tuning(handle, dtype, op=None, u=None, p=None, batch=False):
    """Send picks to a solver (or batch) handle."""
    L = _bind()
    t = Tuning()
    t.xcd_chunk = -1
    if op is not None:
        t.lanes_u, t.lanes_p, t.xcd_chunk = int(op[0]), int(op[1]), int(op[2])
    for table, rows in ((t.u_lpr, u), (t.p_lpr, p)):
        for l, r in enumerate(rows or []):
            for j in range(3):
                table[l][j] = int(r[j])
    fn = L.mpet_batch_set_tuning if batch else L.mpet_solver_set_tuning
    check(fn(handle, dtype, C.byref(t)))
    return t
