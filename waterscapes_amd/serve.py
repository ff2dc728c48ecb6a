"""Serving loop: MPET solve requests admitted by HBM footprint and stepped
by continuous batching on one GPU.

A request is one MPETTotalPressureSolver problem and a number of implicit
time steps.  The server keeps every admitted instance's device state
(assembled operator, Dirichlet data, AMG hierarchies, time-stepping state)
resident until the request completes.  At every scheduling tick it

  1. admits waiting requests in arrival order while their estimated HBM
     footprint fits the budget (a request larger than the whole budget is
     refused with MPET_ERR_BUDGET).  An admitted request's set-up (mesh and
     DOF maps, assembly, host AMG setup, uploads) runs on a pool of host
     threads, each on a stream of its own, while the ticks go on; it joins
     the batch at the first tick after its set-up finished (setup_workers =
     0: set up synchronously inside the tick);
  2. re-packs the batch when requests were admitted (no assembly or AMG
     setup is repeated: each instance's device state is reused);
  3. runs one batched step of the unfinished instances the tick policy
     picks (requests that arrived while others were running join
     mid-flight: their steps overlap on the device with the others');
     "all" steps every unfinished instance; "srpt" (the default) steps
     those with the least remaining work (steps left x DOFs) first, as many
     as make up tick_dofs unknowns together (at least one), so a short
     request is not held behind long ones at every tick while the device
     still gets a full batch; instances not picked, and finished ones, stay
     packed but inactive, so neither costs a re-pack;
  4. retires finished requests (final state copied to the host); their
     memory is released when the batch is next re-packed — at the next
     admission, when half of the packed instances have finished, or when a
     waiting request needs the memory.

The reference has no serving layer (it steps one problem per
MPETTotalPressureSolver, mpet/mpettotalpressuresolver.py:456-502); each
request's steps are that solver's steps.
"""
from __future__ import annotations

import concurrent.futures as cf
import gc
import threading
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .batch import MPETBatch, batch_engine

ERR_BUDGET = 6


class _null:
    """A context manager that does nothing (synchronous set-up: the current
    stream)."""
    def __enter__(self):
        return None

    def __exit__(self, *a):
        return False


class BudgetError(RuntimeError):
    """A request whose footprint exceeds the whole HBM budget."""
    code = ERR_BUDGET


@dataclass
class Request:
    """make_solver() -> MPETTotalPressureSolver with its initial state set
    (called at admission, so waiting requests hold no device memory)."""
    make_solver: object
    steps: int
    arrival: float = 0.0
    rid: object = None
    footprint: int = 0
    # filled by the server
    solver: object = None
    admitted: float = None
    done: float = None
    steps_done: int = 0
    ndofs: int = 0
    result: np.ndarray = None
    niter: list = field(default_factory=list)
    refused: bool = False
    _fp: object = None                 # footprint estimate running on a set-up thread


def footprint_bytes(space, precision="mixed", robin=False):
    """Estimated HBM bytes of one resident instance: FP64 primitives, the
    packed FP64 / FP32 operator copies, the AMG hierarchies with their work
    vectors, and the step / MINRES vectors.  From the node patterns only (no
    device allocation)."""
    dm = space.dofmap
    nz = {k: int(dm.pattern(k)[1].size) for k in ("22", "21", "12", "11")}
    A, n, nP2, nP1 = space.A, space.ndofs, space.nP2, space.nP1
    F = A + 1
    op_vals = 9 * nz["22"] + 3 * nz["21"] + 3 * nz["12"] + (1 + A) * nz["11"]
    idx = 4 * (nz["22"] + nz["21"] + nz["12"] + nz["11"] + nP2 + nP1 + 2)
    prim = 8 * (op_vals + nz["22"] + 3 * nz["12"] + (A * nz["11"] if robin else 0))
    s = 4 if precision == "mixed" else 8
    ops = 8 * op_vals + (s * op_vals if precision == "mixed" else 0) + 2 * idx
    # hierarchy: level 0 (Laplacian values, its pattern is shared) plus the
    # coarse levels (bounded by ~half of level 0) and 7 work vectors per row
    amg = s * nz["22"] + (s + 4) * (nz["22"] + F * nz["11"]) // 2 \
        + 7 * s * (3 * nP2 + F * nP1) * 2
    vecs = 9 * s * n + 13 * 8 * n
    return int(prim + ops + amg + vecs)


class MPETServer:
    """Continuous-batching server over one GPU (see module docstring)."""

    def __init__(self, hbm_budget=200e9, precision="mixed", outer_tol=3e-11, inner_rtol=1e-6,
                 rtol=1e-10, max_instances=512, clock=time.perf_counter, setup_workers=2,
                 policy="srpt", tick_dofs=6e6):
        if not torch.cuda.is_available():
            raise RuntimeError("MPETServer needs a GPU (no CPU fallback)")
        self.budget = float(hbm_budget)
        self.precision = precision
        self.tol = dict(outer_tol=outer_tol, inner_rtol=inner_rtol, rtol=rtol)
        self.max_instances = max_instances
        self.clock = clock
        self.waiting = []          # arrival order
        self.preparing = []        # (request, future): admitted, set-up running
        self.active = []
        self.setup_workers = int(setup_workers)
        if policy not in ("all", "srpt"):
            raise ValueError("policy must be 'all' or 'srpt'")
        self.policy = policy
        self.tick_dofs = float(tick_dofs)
        self._pool = None
        self._local = threading.local()
        self.finished = []
        self.used = 0
        self.batch = None
        self.batch_reqs = []       # requests packed in self.batch, in order
        self.pending_free = 0      # footprint of retired requests still packed
        self.t0 = clock()
        self.steps_run = 0
        self.ticks = 0

    # -- requests ---------------------------------------------------------------
    def submit(self, req: Request):
        self.waiting.append(req)

    def _on_stream(self, fn, req):
        """fn(req) on a set-up thread's own stream, synchronised before the
        instance is used by the batch (synchronous set-up: the current
        stream)."""
        st = None
        if self.setup_workers > 0:
            if not hasattr(self._local, "stream"):
                self._local.stream = torch.cuda.Stream()
            st = self._local.stream
        with torch.cuda.stream(st) if st is not None else _null():
            out = fn(req)
        if st is not None:
            st.synchronize()
        return out

    def _setup(self, req):
        """The request's resident device state (assembly, AMG setup)."""
        def work(r):
            if r.solver is None:
                r.solver = r.make_solver()
            batch_engine(r.solver, self.precision)
            r.ndofs = int(r.solver.up.data.numel())
            return r
        return self._on_stream(work, req)

    def _footprint(self, req):
        """The request's solver (host mesh and DOF maps) and its estimated
        HBM footprint."""
        def work(r):
            if r.footprint == 0:
                if r.solver is None:
                    r.solver = r.make_solver()
                prob = r.solver.problem
                robin = any(not getattr(b, "is_zero", lambda: False)() for b in prob.beta)
                r.footprint = footprint_bytes(r.solver.up_.function_space(), self.precision,
                                              robin)
            return r.footprint
        return self._on_stream(work, req)

    def _submit(self, fn, req):
        if self._pool is None:
            self._pool = cf.ThreadPoolExecutor(max_workers=self.setup_workers)
        return self._pool.submit(fn, req)

    def _footprint_ready(self, req, now):
        """True when the head request's footprint is known; with set-up
        threads, its estimate (and those of the next arrived requests) is
        computed off the scheduling thread."""
        if req.footprint:
            return True
        if self.setup_workers <= 0:
            self._footprint(req)
            return True
        for r in self.waiting[:2 * self.setup_workers]:
            if r.arrival <= now and r.footprint == 0 and r._fp is None:
                r._fp = self._submit(self._footprint, r)
        if req._fp is None or not req._fp.done():
            return False
        req._fp.result()
        return True

    def _admit(self, now):
        changed = False
        # set-ups that finished join the active set
        still = []
        for req, fut in self.preparing:
            if fut.done():
                fut.result()                   # re-raises a set-up failure
                req.admitted = now
                self.active.append(req)
                changed = True
            else:
                still.append((req, fut))
        self.preparing = still
        while self.waiting and len(self.active) + len(self.preparing) < self.max_instances:
            req = self.waiting[0]
            if req.arrival > now or not self._footprint_ready(req, now):
                break
            if req.footprint > self.budget:
                self.waiting.pop(0)
                req.refused = True
                req.solver = None
                req.done = now
                self.finished.append(req)
                continue
            if self.used + req.footprint > self.budget and self.pending_free:
                self._drop_batch()         # release retired instances first
            if self.used + req.footprint > self.budget:
                break                      # wait until memory frees up
            self.waiting.pop(0)
            self.used += req.footprint
            if self.setup_workers > 0:
                self.preparing.append((req, self._submit(self._setup, req)))
            else:
                self._setup(req)           # resident device state
                req.admitted = now
                self.active.append(req)
                changed = True
        return changed

    def busy(self):
        """Requests waiting, being set up or running."""
        return bool(self.waiting or self.preparing or self.active)

    def _drop_batch(self):
        """Close the batch and release the memory of retired instances."""
        if self.batch is not None:
            self.batch.close()
            self.batch = None
        self.batch_reqs = []
        self.used -= self.pending_free
        self.pending_free = 0
        gc.collect()

    def _repack(self):
        self._drop_batch()
        if self.active:
            self.batch = MPETBatch([r.solver for r in self.active], precision=self.precision,
                                   **self.tol)
            self.batch_reqs = list(self.active)

    def _retire(self, now):
        keep = []
        for r in self.active:
            if r.steps_done >= r.steps:
                r.result = r.solver.up.vector().get_local().copy()
                r.done = now
                self.pending_free += r.footprint
                r.solver = None
                self.finished.append(r)
            else:
                keep.append(r)
        changed = len(keep) != len(self.active)
        self.active = keep
        if changed and (not keep or 2 * (len(self.batch_reqs) - len(keep))
                        >= len(self.batch_reqs)):
            self._drop_batch()
        return changed

    # -- scheduling ---------------------------------------------------------------
    def tick(self):
        """Admit, (re)pack, one batched step, retire.  Returns the number of
        instances stepped."""
        now = self.clock() - self.t0
        changed = self._admit(now)
        if not self.active:
            return 0
        if changed or self.batch is None:
            self._repack()
        b = self.batch
        act = self._pick([1 if r.steps_done < r.steps else 0 for r in self.batch_reqs])
        stats = b.step(act)
        b.commit(None if act.all() else act)
        for k, r in enumerate(self.batch_reqs):
            if act[k]:
                r.steps_done += 1
                r.niter.append(int(stats[k].inner_iterations))
        stepped = int(act.sum())
        self.steps_run += stepped
        self.ticks += 1
        torch.cuda.synchronize()
        self._retire(self.clock() - self.t0)
        return stepped

    def _pick(self, unfinished):
        """Instances (0/1 per packed request) the tick steps under the policy."""
        act = np.array(unfinished, dtype=np.int32)
        if self.policy == "all" or act.sum() <= 1:
            return act
        reqs = self.batch_reqs
        cand = sorted((k for k in range(len(reqs)) if act[k]),
                      key=lambda k: ((reqs[k].steps - reqs[k].steps_done) * reqs[k].ndofs,
                                     reqs[k].arrival, k))
        sel = np.zeros_like(act)
        tot = 0
        for k in cand:
            sel[k] = 1
            tot += reqs[k].ndofs
            if tot >= self.tick_dofs:
                break
        return sel

    def run(self, until_idle=True, idle_sleep=1e-3):
        """Tick until every submitted request has finished."""
        while self.busy():
            if not self.tick():
                head = self.waiting[0] if self.waiting else None
                if self.preparing or (head is not None and head._fp is not None
                                      and not head._fp.done()):
                    time.sleep(idle_sleep)         # a set-up or estimate is running
                elif self.waiting and self.waiting[0].arrival > self.clock() - self.t0:
                    time.sleep(min(idle_sleep, self.waiting[0].arrival -
                                   (self.clock() - self.t0)))
                elif self.waiting and not self.active:
                    # the head request does not fit although nothing runs:
                    # only possible when it exceeds the budget (refused above)
                    raise BudgetError("request cannot be admitted")
        self._drop_batch()
        self.close()
        return self.finished

    def close(self):
        """Stop the set-up threads (after the running set-ups finish)."""
        if self._pool is not None:
            self._pool.shutdown(wait=True)
            self._pool = None

    def latencies(self):
        lat = [r.done - r.arrival for r in self.finished if not r.refused]
        return np.array(lat)
