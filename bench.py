"""Benchmark: implicit-Euler steps/s of the total-pressure MPET step on one
MI355X, mixed precision vs plain FP64, on the 48^3 x 6 tetrahedra, A = 4
network workload (663,552 cells, 3,326,264 DOFs):

  E = 1500, nu = 0.479, alpha_j = 0.25, c_j = 1e-5, K_j per cell log-normal
  (medians 1e-6, 1e-7, 1e-6, 1e-5, sigma 0.5, fixed seed), S_{j,j+-1} = 1e-3;
  u = 0 on z = 0; p_j = p0_j + 5 sin(2 pi t) on z = 1 (p0 = 8, 4, 2, 1);
  theta = 1, dt = 0.0125.
  FP64 path: MINRES + block AMG to rtol 1e-10 (scalar CSR FP64 operator).
  Mixed path: FP32 MINRES / AMG inside FP64 iterative refinement.
  Accuracy: per step and per field, relative L2 error of both paths against
  an FP64 reference solve of the same steps to rtol 1e-13 (untimed).

Prints one JSON line (rank 0).  `value` = mixed-precision steps/s.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

P0 = (8.0, 4.0, 2.0, 1.0, 0.5, 0.25, 0.125)
MEDIANS = (1e-6, 1e-7, 1e-6, 1e-5, 1e-6, 1e-6, 1e-6)


def workload_params(A):
    S = [[1e-3 if abs(i - j) == 1 else 0.0 for j in range(A)] for i in range(A)]
    return dict(J=A, E=1500.0, nu=0.479, alpha=(0.25,) * A, c=(1e-5,) * A,
                K=MEDIANS[:A], S=S)


def lognormal_cells(nC, A, seed):
    rng = np.random.default_rng(seed)
    med = np.asarray(MEDIANS[:A])
    return med[None, :] * np.exp(0.5 * rng.standard_normal((nC, A)))


def make_solver(n, A, precision, seed=2024, **extra):
    from waterscapes_amd import (MPETProblem, MPETTotalPressureSolver, Constant, Expression,
                                 UnitCubeMesh, CompiledSubDomain)
    mesh = UnitCubeMesh(n, n, n)
    time_c = Constant(0.0)
    params = workload_params(A)
    Kc = lognormal_cells(mesh.num_cells(), A, seed)
    params["K"] = tuple(Kc[:, j] for j in range(A))
    problem = MPETProblem(mesh, time_c, params=params)
    problem.p_bar = [Expression(f"{P0[j]} + 5.0*sin(2.0*M_PI*t)", t=time_c, degree=1)
                     for j in range(A)]
    CompiledSubDomain("on_boundary && near(x[2], 0.0)").mark(problem.momentum_boundary_markers, 0)
    for j in range(A):
        CompiledSubDomain("on_boundary && near(x[2], 1.0)").mark(
            problem.continuity_boundary_markers[j], 0)
    sp = dict(dt=0.0125, theta=1.0, T=1e9, direct_solver=False, precision=precision,
              krylov_rtol=1e-10)
    sp.update(extra)
    solver = MPETTotalPressureSolver(problem, sp)
    return solver


def run_steps(solver, W, K, keep=False):
    """W untimed + K timed steps; returns (seconds, per-step seconds, states,
    niter, outer)."""
    import torch
    gen = solver.solve()
    for _ in range(W):
        next(gen)
    torch.cuda.synchronize()
    states, lat = [], []
    t0 = time.perf_counter()
    for _ in range(K):
        s0 = time.perf_counter()
        up, t = next(gen)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - s0)
        if keep:
            states.append(up.data.clone())
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    mon = solver.solver_monitor
    return total, lat, states, mon["niter"][W:], mon["outer"][W:]


def op_bytes(sys_, dtype):
    """Algorithmic HBM bytes of one structured operator apply y = A x:
    values + column indices + row pointers of the four node patterns, x
    read once, y written once."""
    s = 4 if dtype == 1 else 8
    nz = sys_.nnz
    A = sys_.A
    vals = (9 * nz["22"] + 3 * nz["21"] + 3 * nz["12"] + (1 + A) * nz["11"]) * s
    idx = 4 * (nz["22"] + nz["21"] + nz["12"] + nz["11"]) + 4 * 2 * (sys_.nP2 + 1 + sys_.nP1 + 1)
    vec = 2 * sys_.n * s
    return vals + idx + vec


def cell_op_bytes(sys_, s=4):
    """Algorithmic HBM bytes of one cell-wise operator apply (s = 4: FP32,
    8: FP64 with FP64 geometry, partials, P1 x P1 values and vectors)
    (csrc/cellwise.h), each array once: the cell kernel reads the node
    tables (14 ints per cell), the geometry (10 floats), x and the block
    partial tables (block and slot offsets, 2-byte local slots), and writes
    the partials (3 floats per P2 partial, 1 per P1 partial, block-major);
    the row kernel reads the node offsets, the entries (one partial index
    each), the partials, the P1 x P1 block (indices, M1, A stiffness
    arrays) and the pressure part of x, and writes y."""
    cw = sys_.cell_op()
    nC = cw["nC"]
    n, F, A = sys_.n, sys_.F, sys_.A
    nz11 = sys_.nnz["11"]
    nb = (nC + sys_.CELLS_PER_BLOCK - 1) // sys_.CELLS_PER_BLOCK
    p2, p1 = cw["p2"]["npart"], cw["p1"]["npart"]
    tables = (nb + 1) * 4 * 2 + (p2 + 1) * 4 + 10 * nC * 2 + (p1 + 1) * 4 + 4 * nC * 2
    cells = nC * (14 * 4 + 10 * s) + n * s + tables + (3 * p2 + p1) * s
    rows = ((sys_.nP2 + 1) + 2 * (sys_.nP1 + 1)) * 4 + (p2 + p1) * 4 + (3 * p2 + p1) * s \
        + nz11 * (4 + s + s * A) + F * sys_.nP1 * s + n * s
    return cells + rows


def csr_bytes(sys_):
    rp, ci, v = sys_.csr
    nnz = v.numel()
    return nnz * (8 + 4) + (sys_.n + 1) * 8 + 2 * sys_.n * 8


def _mat_bytes(nnz, rows, s, compact=False):
    """Values + column indices + row pointers of one CSR level matrix (the
    compact FP32 finest levels: half values and 16-bit column offsets)."""
    return nnz * (4 if compact else s + 4) + (rows + 1) * 4


def vcycle_bytes(levels, amgp, NC, s, compact=(), scaled=False):
    """Algorithmic HBM bytes of one V-cycle (amg_cycle.h: vcycle / smooth /
    spmm modes), each array read once per launch: matrices, the gathered
    vector once, the epilogue vectors; levels = the host hierarchy dicts,
    compact = the levels stored as half values with 16-bit offsets.  The
    Chebyshev direction is not read by the first step from x and not written
    by a sequence's last step."""
    nl = len(levels)
    tot = 0

    def spmm(mode, nnz, rows, cols, cmp=False, dvec=2):
        mat = _mat_bytes(nnz, rows, s, cmp) + NC * cols * s
        epi = {0: 1, 1: 2, 2: 2, 3: 2 + dvec, 4: 3}[mode] * NC * rows * s
        if mode in (3, 4):
            epi += rows * s                      # D^-1
        return mat + epi

    def level(l, entered):
        nonlocal tot
        lev = levels[l]
        n = lev["n"]
        if l == nl - 1:
            tot += n * n * s + 2 * NC * n * s      # dense coarsest solve
            return
        nnz = len(lev["A_ci"])
        cmp = l in compact
        deg = amgp.degree(l)
        if not entered:                            # first Chebyshev step from 0
            tot += ((3 if deg > 1 else 2) + (1 if (scaled and l == 0) else 0)) * NC * n * s \
                + n * s
        if deg > 1:                                # the last one writes no direction
            tot += (deg - 2) * spmm(3, nnz, n, n, cmp) + spmm(3, nnz, n, n, cmp, 1)
        tot += spmm(1, nnz, n, n, cmp)             # residual
        nc = lev["nc"]
        nnzP = len(lev["P_ci"])
        tot += spmm(4 if l + 1 < nl - 1 else 0, nnzP, nc, n)   # restriction
        level(l + 1, l + 1 < nl - 1)
        tot += spmm(2, nnzP, n, nc)                # prolongation
        # post-smoothing from x: the first step reads no direction
        if deg == 1:
            tot += spmm(3, nnz, n, n, cmp, 0)
        else:
            tot += (spmm(3, nnz, n, n, cmp, 1) + (deg - 2) * spmm(3, nnz, n, n, cmp)
                    + spmm(3, nnz, n, n, cmp, 1))
    level(0, False)
    return tot


def _compact_u(kr, mixed):
    return (0,) if mixed and getattr(kr, "compact", None) is not None else ()


def _compact_p(kr, mixed):
    return tuple(getattr(kr, "pcompact_levels", ())) if mixed else ()


def iteration_bytes(eng, mixed):
    """Algorithmic HBM bytes of one MINRES iteration of the engine's solve
    (krylov.hip minres_iter): K1 (operator apply with the fused MINRES step:
    + r1 read; v is not stored), the block preconditioner (displacement and
    pressure V-cycles, the pressure gather / scatter) and the BLAS-1 work
    (the K2 update y, r_old -> r made inside the preconditioner's first
    kernels, 3n, less the n those kernels no longer read; dot 2n; k5 6n).
    Returns (bytes, parts)."""
    sysm = eng.system
    kr = eng.krylov
    s = 4 if mixed else 8
    n = sysm.n
    if mixed and getattr(kr, "cell_desc", None) is not None:
        k1 = cell_op_bytes(sysm) + n * s
    elif mixed or kr.csr_desc is None:
        k1 = op_bytes(sysm, 1 if mixed else 0) + n * s
    else:
        k1 = csr_bytes(sysm) + n * s
    vu = vcycle_bytes(kr.levels, kr.cfg.amg, 3, s, compact=_compact_u(kr, mixed), scaled=True)
    vp = vcycle_bytes(kr.plevels, kr.cfg.pamg, 1, s,
                      compact=_compact_p(kr, mixed)) + 4 * sysm.F * sysm.nP1 * s
    blas = 10 * n * s
    parts = {"k1": int(k1), "vcycle_u": int(vu), "vcycle_p": int(vp), "blas1": int(blas)}
    return int(k1 + vu + vp + blas), parts


def step_bytes(eng, mixed, inner, outer):
    """Algorithmic bytes of one time step: inner iterations, the FP64
    residuals (through the cell-wise FP64 operator on the mixed path) and
    MINRES starts of the solve (two for the norm of the
    lifted right-hand side and the first residual, one more per refinement
    step) and the right-hand-side update (P1-pattern kernel)."""
    sysm = eng.system
    it_b, _ = iteration_bytes(eng, mixed)
    s = 4 if mixed else 8
    n = sysm.n
    kr = eng.krylov
    if mixed and getattr(kr, "cell_desc64", None) is not None:
        op64 = cell_op_bytes(sysm, 8)
    else:
        op64 = op_bytes(sysm, 0) if (mixed or kr.csr_desc is None) else csr_bytes(sysm)
    pc = vcycle_bytes(kr.levels, kr.cfg.amg, 3, s, compact=_compact_u(kr, mixed),
                      scaled=True) + vcycle_bytes(kr.plevels, kr.cfg.pamg, 1, s,
                                                  compact=_compact_p(kr, mixed))
    starts = 2 + (outer if mixed else 0)
    solve_overhead = starts * (op64 + n * 8 + pc + 2 * n * s)
    nz11 = sysm.nnz["11"]
    rhs = nz11 * ((sysm.A + 1) * 8 + 4) + (sysm.nP1 + 1) * 4 + 2 * n * 8
    return int(inner * it_b + solve_overhead + rhs)


def time_kernel(fn, reps=20, warm=3):
    """Average duration (ms) of fn() on the current stream via HIP events."""
    import torch
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        fn()
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def cpu_baseline(n_sample, A, steps, budget_s=30.0):
    """The FP64 oracle (numpy/scipy restatement, direct LU as in the
    reference's default solve_direct) on the same workload family at a
    reduced mesh, timed on this host; extrapolated to the full workload by
    DOF count (linear, which favours the CPU: LU cost grows faster)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import mpet_oracle as O
    from waterscapes_amd.fem import MixedSpace
    from waterscapes_amd.mesh import UnitCubeMesh
    from conftest import oracle_mesh
    mesh = UnitCubeMesh(n_sample, n_sample, n_sample)
    space = MixedSpace(mesh, A)
    m = oracle_mesh(space)
    params = workload_params(A)
    Kc = lognormal_cells(mesh.num_cells(), A, 7)
    mm = np.full(mesh.num_facets(), 2**63 - 1, dtype=np.uint64)
    cm = np.full(mesh.num_facets(), 2**63 - 1, dtype=np.uint64)
    mid = mesh.facet_midpoints()
    mm[mesh.exterior & (np.abs(mid[:, 2]) < 1e-12)] = 0
    cm[mesh.exterior & (np.abs(mid[:, 2] - 1) < 1e-12)] = 0
    pb = [lambda X, t, j=j: np.full(X.shape[0], P0[j] + 5.0 * np.sin(2 * np.pi * t))
          for j in range(A)]
    t0 = time.perf_counter()
    prob = O.OracleProblem(m, params, dict(dt=0.0125, theta=1.0, T=1e9), K_cells=Kc,
                           u_bar=lambda X, t: np.zeros((X.shape[0], 3)), p_bar=pb,
                           momentum_markers=mm, continuity_markers=[cm] * A, refine_steps=0)
    gen = prob.solve_direct(np.zeros(space.ndofs))
    next(gen)                                      # factorisation + first step
    setup = time.perf_counter() - t0
    t1 = time.perf_counter()
    k = 0
    while k < steps and time.perf_counter() - t1 < budget_s:
        next(gen)
        k += 1
    per_step = (time.perf_counter() - t1) / max(k, 1)
    return dict(ndofs=space.ndofs, per_step=per_step, setup=setup, steps=k)


def cpu_baseline_headline(eng, iters_per_step, iters=4, budget_s=40.0):
    """The plain FP64 path of the full workload on the host: textbook MINRES
    (the oracle's) with the same block preconditioner restated in numpy
    (oracle/vcycle.py) on the same matrices (the GPU-assembled FP64 CSR of A
    and the host AMG hierarchies of the engine), scipy sparse products.  A
    bounded sample: `iters` MINRES iterations of the last timed step (plus
    the start, counted as one more), scaled by the GPU FP64 path's mean
    iterations per step at the same tolerance.  Single-threaded sparse
    products (scipy), numpy BLAS-1."""
    import gc
    import scipy.sparse as sp
    from oracle import mpet_oracle as O
    from oracle.vcycle import block_preconditioner
    sysm = eng.system
    kr = eng.krylov
    rp, ci, v = sysm.csr
    n = sysm.n
    A = sp.csr_matrix((v.cpu().numpy(), ci.cpu().numpy(), rp.cpu().numpy()), shape=(n, n))
    dm = sysm.dmask.cpu().numpy().astype(bool)
    rpl, cil = sysm.pat_host["22"]
    lap = sp.csr_matrix((sysm.prim_t["Lap"].cpu().numpy(), cil, rpl), shape=(sysm.nP2,) * 2)
    b = eng.b.cpu().numpy().copy()
    x0 = np.zeros(n)
    x0[dm] = 0.0

    def Aop(x):
        y = A @ x
        y[dm] = 0.0
        return y

    def Minv(r):
        r = r.copy()
        r[dm] = 0.0
        z = block_preconditioner(kr, r, A0=lap)
        z[dm] = 0.0
        return z
    rr = b.copy()
    rr[dm] = 0.0
    t0 = time.perf_counter()
    O.minres(Aop, rr, Minv, x0=x0, rtol=1e-300, maxiter=iters)
    wall = time.perf_counter() - t0
    per_it = wall / (iters + 1)
    del A, lap
    gc.collect()
    return dict(per_iteration=per_it, iters=iters, wall=wall,
                per_step=per_it * iters_per_step)


def uneven_instance(rng, big=False):
    """One instance of the uneven-batch workload family: N ~ U{8..32} (48 for
    a large request), A ~ U{1..4} (4), E log-uniform [1e3, 1e5], nu ~
    U[0.30, 0.49], K_j log-uniform [1e-6, 1e-2], S_ij log-uniform
    [1e-4, 1e-2] (symmetric), alpha_j = 0.25, c_j = 1e-5; u = 0 on z = 0,
    p_j = sin(2 pi t) on z = 1; theta = 1, dt = 0.0125."""
    N = 48 if big else int(rng.integers(8, 33))
    A = 4 if big else int(rng.integers(1, 5))
    E = float(10 ** rng.uniform(3, 5))
    nu = float(rng.uniform(0.30, 0.49))
    K = tuple(float(k) for k in 10 ** rng.uniform(-6, -2, A))
    S = np.zeros((A, A))
    for i in range(A):
        for j in range(i + 1, A):
            S[i, j] = S[j, i] = 10 ** rng.uniform(-4, -2)
    params = dict(J=A, E=E, nu=nu, alpha=(0.25,) * A, c=(1e-5,) * A, K=K,
                  S=[list(r) for r in S])
    return N, params


def make_uneven_solver(N, params, precision="mixed", **extra):
    from waterscapes_amd import (MPETProblem, MPETTotalPressureSolver, Constant, Expression,
                                 UnitCubeMesh, CompiledSubDomain)
    A = params["J"]
    mesh = UnitCubeMesh(N, N, N)
    time_c = Constant(0.0)
    problem = MPETProblem(mesh, time_c, params=params)
    problem.p_bar = [Expression("sin(2.0*M_PI*t)", t=time_c, degree=1) for _ in range(A)]
    CompiledSubDomain("on_boundary && near(x[2], 0.0)").mark(problem.momentum_boundary_markers, 0)
    for j in range(A):
        CompiledSubDomain("on_boundary && near(x[2], 1.0)").mark(
            problem.continuity_boundary_markers[j], 0)
    sp = dict(dt=0.0125, theta=1.0, T=1e9, direct_solver=False, precision=precision,
              krylov_rtol=1e-10)
    sp.update(extra)
    return MPETTotalPressureSolver(problem, sp)


def main_batch(args):
    """Uneven batch: `--instances` instances stepped together (mixed), vs the
    same instances one at a time with the plain FP64 GPU path (timed on a
    random sample of `--sample` instances, scaled to the batch)."""
    print(json.dumps(run_batch(args, args.instances, args.steps, args.warmup, args.sample)))


def run_batch(args, instances, steps, warmup, sample):
    """The uneven-batch measurement of main_batch; returns its JSON object."""
    import torch
    from waterscapes_amd.batch import MPETBatch
    rng = np.random.default_rng(args.seed)
    specs = [uneven_instance(rng) for _ in range(instances)]
    t0 = time.perf_counter()
    solvers = [make_uneven_solver(N, p) for N, p in specs]
    batch = MPETBatch(solvers, precision="mixed", outer_tol=args.outer_tol,
                      inner_rtol=args.inner_rtol, autotune=True)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    ndofs = sum(s.up.data.numel() for s in solvers)
    print(f"[bench] batch of {len(solvers)} ({ndofs} DOFs) set up in {setup:.1f}s",
          file=sys.stderr, flush=True)
    for _ in range(warmup):
        batch.step()
        batch.commit()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    iters = []
    for _ in range(steps):
        st = batch.step()
        batch.commit()
        iters.append(max(int(st[k].inner_iterations) for k in range(batch.K)))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    inst_steps = batch.K * steps
    mem = torch.cuda.max_memory_allocated()
    tuning = batch.tuning
    batch.close()
    del batch, solvers
    import gc
    gc.collect()
    torch.cuda.empty_cache()
    # plain FP64 per-instance path on a sample
    idx = rng.choice(len(specs), size=min(sample, len(specs)), replace=False)
    seq_time, seq_steps = 0.0, 0
    for i in idx:
        N, p = specs[i]
        s = make_uneven_solver(N, p, precision="fp64")
        s.engine()
        tot = run_steps(s, warmup, steps)[0]
        seq_time += tot
        seq_steps += steps
        del s
        gc.collect()
    seq_rate = seq_steps / seq_time
    rate = inst_steps / wall
    out = {
        "metric": "uneven batch: instance-steps/s (mixed, batched) vs plain FP64 one at a time",
        "value": round(rate, 4), "unit": "instance-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": round(1e3 * wall / steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32-in-fp64", "data": "synthetic",
        "config": {"workload": f"{len(specs)} uneven MPET instances (N 8..32, A 1..4), "
                               f"theta=1, dt=0.0125", "dofs": int(ndofs),
                   "seed": args.seed, "outer_tol": args.outer_tol,
                   "inner_rtol": args.inner_rtol},
        "dof_steps_per_s": round(rate * ndofs / len(specs), 1),
        "max_inner_iterations_per_step": iters,
        "setup_s": round(setup, 2), "max_hbm_bytes": int(mem),
        "autotune": tuning,
        "fp64_one_at_a_time": {"instance_steps_per_s": round(seq_rate, 4),
                               "sample": [int(i) for i in idx]},
        "speedup_vs_fp64_one_at_a_time": round(rate / seq_rate, 3),
    }
    return out


def serve_specs(seed, nreq, steps_scale):
    """Request specs of the serving stream: the uneven family with one
    request in 20 at N = 48, A = 4, 10..100 steps each (scaled), and the
    unit-rate exponential gaps of the Poisson arrivals."""
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(nreq):
        big = (k % 20) == 19
        N, p = uneven_instance(rng, big)
        steps = max(1, int(round(rng.integers(10, 101) * steps_scale)))
        specs.append((N, p, steps))
    gaps = rng.exponential(1.0, nreq)
    return specs, np.cumsum(gaps)


def fp64_service_times(specs):
    """Seconds each request takes alone on the plain FP64 GPU path (set-up
    plus its steps; scalar CSR operator, guess up_, rtol 1e-10)."""
    import gc
    import torch
    out = []
    for N, p, steps in specs:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s = make_uneven_solver(N, p, precision="fp64")
        gen = s.solve()
        for _ in range(steps):
            next(gen)
        torch.cuda.synchronize()
        out.append(time.perf_counter() - t0)
        print(f"[serve] FP64 alone: request {len(out)}/{len(specs)} (N={N}, {steps} steps) "
              f"{out[-1]:.2f} s", file=sys.stderr, flush=True)
        del gen, s
        gc.collect()
    return out


def main_serve(args):
    """Serving sweep: for each rate in `--rates`, the same `--requests`
    requests (uneven family, 1 in 20 at 48^3 A = 4, 10..100 steps) arrive as
    a Poisson stream at that rate and are served by continuous batching
    (admission by HBM footprint, mixed precision); p50/p99 latency (arrival
    to last step) and instance-steps/s.  Beside each: the plain FP64 path
    serving the same arrivals one request at a time (FIFO; each request's
    service time measured alone once, the queue computed from the
    arrivals).  One JSON line per rate."""
    import torch
    from waterscapes_amd.serve import MPETServer, Request
    specs, base = serve_specs(args.seed, args.requests, args.steps_scale)
    rates = [float(r) for r in str(args.rates).split(",")]
    print(f"[serve] {len(specs)} requests, {sum(s[2] for s in specs)} instance-steps, "
          f"{sum(1 for s in specs if s[0] == 48)} at 48^3", file=sys.stderr, flush=True)
    service = None
    if not args.skip_fp64:
        service = fp64_service_times(specs)
        print(f"[serve] FP64 one at a time: {sum(service):.1f} s of service", file=sys.stderr,
              flush=True)
    for rate in rates:
        arrivals = base / rate
        reqs = [Request(make_solver=lambda N=N, p=p: make_uneven_solver(N, p), steps=st,
                        arrival=float(a), rid=k)
                for k, ((N, p, st), a) in enumerate(zip(specs, arrivals))]
        torch.cuda.reset_peak_memory_stats()
        srv = MPETServer(hbm_budget=args.budget * 1e9, precision="mixed",
                         outer_tol=args.outer_tol, inner_rtol=args.inner_rtol,
                         policy=args.policy, tick_dofs=args.tick_dofs)
        for r in reqs:
            srv.submit(r)
        t0 = time.perf_counter()
        last = t0
        while srv.busy():
            if not srv.tick():
                time.sleep(1e-3)
            if time.perf_counter() - last > 10:
                last = time.perf_counter()
                print(f"[serve] rate {rate}: t={last - t0:.0f}s active {len(srv.active)} "
                      f"waiting {len(srv.waiting)} finished {len(srv.finished)}",
                      file=sys.stderr, flush=True)
        srv._drop_batch()
        srv.close()
        wall = time.perf_counter() - t0
        lat = srv.latencies()
        done = [r for r in srv.finished if not r.refused]
        steps = sum(r.steps for r in done)
        out = {
            "metric": "served MPET requests: p50/p99 latency, aggregate instance-steps/s",
            "value": round(steps / wall, 4), "unit": "instance-steps/s", "n_gpus": 1,
            "steps": int(steps), "warmup": 0,
            "ms_per_step": round(1e3 * wall / max(srv.ticks, 1), 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "fp32-in-fp64", "data": "synthetic",
            "config": {"workload": f"Poisson arrivals at {rate} req/s of {len(specs)} requests, "
                                   f"uneven family + 1 in 20 at 48^3 A=4, budget "
                                   f"{args.budget} GB", "rate": rate, "requests": len(specs),
                       "seed": args.seed, "policy": args.policy,
                       "tick_dofs": args.tick_dofs},
            "latency_s": {"p50": round(float(np.percentile(lat, 50)), 3),
                          "p99": round(float(np.percentile(lat, 99)), 3)} if lat.size else None,
            "refused": sum(r.refused for r in srv.finished),
            "ticks": srv.ticks, "wall_s": round(wall, 2),
            "max_hbm_bytes": int(torch.cuda.max_memory_allocated()),
        }
        if service is not None:
            fin, lat64 = 0.0, []
            for a, sv in zip(arrivals, service):
                fin = max(fin, a) + sv
                lat64.append(fin - a)
            lat64 = np.array(lat64)
            span = fin - float(arrivals[0])
            out["fp64_one_at_a_time"] = {
                "instance_steps_per_s": round(steps / max(fin, 1e-9), 4),
                "latency_s": {"p50": round(float(np.percentile(lat64, 50)), 3),
                              "p99": round(float(np.percentile(lat64, 99)), 3)},
                "makespan_s": round(fin, 2),
                "note": "FIFO over the same arrivals; per-request service times (set-up + "
                        "steps, plain FP64 path) measured alone once"}
            out["speedup_vs_fp64_one_at_a_time"] = round((steps / wall) / (steps / fin), 3)
        print(json.dumps(out), flush=True)


# two-field MPETSolver workload: the reference's own CPU-runnable material
# set (cfg1: E 1e4, nu 0.35, two networks) on a larger cube, with the
# cfg2-style forcing (u = 0 at z = 0, oscillating network pressures at z = 1)
TF_PARAMS = dict(J=2, E=1e4, nu=0.35, alpha=(0.5, 0.5), c=(1e-4, 1e-4), K=(1e-3, 1e-4),
                 S=((0.0, 1e-2), (1e-2, 0.0)))


def make_two_field_solver(n, dt=0.0125, precision="fp64"):
    from waterscapes_amd import (MPETProblem, MPETSolver, Constant, Expression, UnitCubeMesh,
                                 CompiledSubDomain)
    mesh = UnitCubeMesh(n, n, n)
    time_c = Constant(0.0)
    problem = MPETProblem(mesh, time_c, params=dict(TF_PARAMS))
    problem.p_bar = [Expression(f"{P0[j]} + 5.0*sin(2.0*M_PI*t)", t=time_c, degree=1)
                     for j in range(2)]
    CompiledSubDomain("on_boundary && near(x[2], 0.0)").mark(problem.momentum_boundary_markers, 0)
    for j in range(2):
        CompiledSubDomain("on_boundary && near(x[2], 1.0)").mark(
            problem.continuity_boundary_markers[j], 0)
    return MPETSolver(problem, dict(dt=dt, theta=1.0, T=1e9, precision=precision))


def two_field_cpu_baseline(n_sample, steps, budget_s=20.0):
    """The oracle's two-field restatement (SuperLU, as the reference's MUMPS
    direct path) on the same workload at a reduced mesh, per step and with
    the factorisation amortised over 80 steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import mpet_oracle as O
    from waterscapes_amd.fem import MixedSpace
    from waterscapes_amd.mesh import UnitCubeMesh
    from conftest import oracle_mesh
    mesh = UnitCubeMesh(n_sample, n_sample, n_sample)
    space = MixedSpace(mesh, 2, total_pressure=False)
    m = oracle_mesh(space)
    mm = np.full(mesh.num_facets(), 2**63 - 1, dtype=np.uint64)
    cm = np.full(mesh.num_facets(), 2**63 - 1, dtype=np.uint64)
    mid = mesh.facet_midpoints()
    mm[mesh.exterior & (np.abs(mid[:, 2]) < 1e-12)] = 0
    cm[mesh.exterior & (np.abs(mid[:, 2] - 1) < 1e-12)] = 0
    pb = [lambda X, t, j=j: np.full(X.shape[0], P0[j] + 5.0 * np.sin(2 * np.pi * t))
          for j in range(2)]
    t0 = time.perf_counter()
    prob = O.OracleProblem(m, TF_PARAMS, dict(dt=0.0125, theta=1.0, T=1e9),
                           u_bar=lambda X, t: np.zeros((X.shape[0], 3)), p_bar=pb,
                           momentum_markers=mm, continuity_markers=[cm] * 2,
                           formulation="two_field", refine_steps=0)
    gen = prob.solve_direct(np.zeros(space.ndofs))
    next(gen)
    setup = time.perf_counter() - t0
    t1 = time.perf_counter()
    k = 0
    while k < steps and time.perf_counter() - t1 < budget_s:
        next(gen)
        k += 1
    per_step = (time.perf_counter() - t1) / max(k, 1)
    return dict(ndofs=space.ndofs, per_step=per_step, setup=setup, steps=k)


def main_two_field(args):
    """Two-field MPETSolver (mpet/mpetsolver.py) steps/s: the reference's
    direct path (MUMPS LU) realised as FP64 MINRES to 1e-12 on the device
    CSR system with the block AMG preconditioner of prec."""
    import torch
    from waterscapes_amd import _lib
    n = args.tf_n
    K, W = min(args.steps, 10), args.warmup
    t_setup = time.perf_counter()
    s = make_two_field_solver(n, precision=args.tf_precision)
    eng = s.engine()
    torch.cuda.synchronize()
    setup = time.perf_counter() - t_setup
    gen = s.solve()
    for _ in range(W):
        next(gen)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        next(gen)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    its = s.solver_monitor["niter"][W:]
    sysd = eng.system
    x = torch.randn(sysd.n, device="cuda", dtype=torch.float64)
    lpr = eng.krylov.csr_desc.lpr
    ms = time_kernel(lambda: sysd.csr_apply(x, lpr=lpr, masked=False), reps=20)
    nbytes = csr_bytes(sysd)
    ach = nbytes / (ms * 1e-3) / 1e9
    mixed = args.tf_precision == "mixed"
    out = {"metric": "two-field MPETSolver steps/s (MINRES to 1e-12 relative in place of LU)",
           "value": round(K / total, 4), "unit": "steps/s", "n_gpus": 1, "steps": K,
           "warmup": W, "ms_per_step": round(1e3 * total / K, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None,
           "dtype": "fp32-in-fp64" if mixed else "f64", "data": "synthetic",
           "config": {"workload": f"two-field MPET, unit cube {n}^3 x 6 tets, A=2 (E 1e4, "
                                  "nu 0.35), theta=1, dt=0.0125", "dofs": int(sysd.n),
                      "precision": args.tf_precision},
           "dof_per_s": round(sysd.n * K / total, 1), "iterations": its,
           "setup_s": round(setup, 2),
           "roofline": {"bound": "hbm", "kernel": f"csr_spmv_kernel<double,{lpr}> (two-field A)",
                        "achieved": round(ach, 1), "peak": 8000.0, "unit": "GB/s",
                        "frac": round(ach / 8000.0, 4), "traffic": None,
                        "bytes_per_launch": int(nbytes), "ms_per_launch": round(ms, 4)}}
    if not args.no_cpu_baseline:
        cb = two_field_cpu_baseline(args.cpu_n, 20)
        per_full = (cb["per_step"] + cb["setup"] / 80.0) * sysd.n / cb["ndofs"]
        out["cpu_baseline"] = {
            "value": round(1.0 / per_full, 6), "unit": "steps/s (extrapolated by DOF count)",
            "cores": 1, "kind": "port",
            "sample": f"two-field FP64 oracle (numpy/scipy, SuperLU as the reference's LU) "
                      f"at {args.cpu_n}^3 ({cb['ndofs']} DOFs), {cb['steps']} steps after "
                      f"factorisation, {cb['per_step']:.3f} s/step, setup {cb['setup']:.1f} s "
                      "amortised over 80 steps"}
    print(json.dumps(out))


def make_nullspace_solver(n, A, seed=2024):
    """The 48^3 A = 4 network workload without the displacement Dirichlet
    condition: traction-free everywhere, rigid motions removed by the
    multipliers of u_has_nullspace (mpettotalpressuresolver.py:179-202,
    285-293); network pressures oscillating on z = 1.  FP64 MINRES to rtol
    1e-10 on the bordered system (iterative path), initial guess
    extrapolated in time (quartic)."""
    from waterscapes_amd import (MPETProblem, MPETTotalPressureSolver, Constant, Expression,
                                 UnitCubeMesh, CompiledSubDomain)
    mesh = UnitCubeMesh(n, n, n)
    time_c = Constant(0.0)
    params = workload_params(A)
    Kc = lognormal_cells(mesh.num_cells(), A, seed)
    params["K"] = tuple(Kc[:, j] for j in range(A))
    problem = MPETProblem(mesh, time_c, params=params)
    problem.u_has_nullspace = True
    problem.p_bar = [Expression(f"{P0[j]} + 5.0*sin(2.0*M_PI*t)", t=time_c, degree=1)
                     for j in range(A)]
    for j in range(A):
        CompiledSubDomain("on_boundary && near(x[2], 1.0)").mark(
            problem.continuity_boundary_markers[j], 0)
    return MPETTotalPressureSolver(problem, dict(dt=0.0125, theta=1.0, T=1e9,
                                                 direct_solver=False, precision="fp64",
                                                 krylov_rtol=1e-10, extrapolate=True))


def nullspace_cpu_baseline(n_sample, A, steps, budget_s=20.0):
    """The oracle's bordered LU (u nullspace) on the same workload at a
    reduced mesh, per step with the factorisation amortised over 80 steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import mpet_oracle as O
    from waterscapes_amd.fem import MixedSpace
    from waterscapes_amd.mesh import UnitCubeMesh
    from conftest import oracle_mesh
    mesh = UnitCubeMesh(n_sample, n_sample, n_sample)
    space = MixedSpace(mesh, A, u_nullspace=True)
    m = oracle_mesh(space)
    params = workload_params(A)
    Kc = lognormal_cells(mesh.num_cells(), A, 7)
    mm = np.full(mesh.num_facets(), 2**63 - 1, dtype=np.uint64)
    cm = np.full(mesh.num_facets(), 2**63 - 1, dtype=np.uint64)
    mid = mesh.facet_midpoints()
    cm[mesh.exterior & (np.abs(mid[:, 2] - 1) < 1e-12)] = 0
    pb = [lambda X, t, j=j: np.full(X.shape[0], P0[j] + 5.0 * np.sin(2 * np.pi * t))
          for j in range(A)]
    t0 = time.perf_counter()
    prob = O.OracleProblem(m, params, dict(dt=0.0125, theta=1.0, T=1e9), K_cells=Kc,
                           u_bar=lambda X, t: np.zeros((X.shape[0], 3)), p_bar=pb,
                           momentum_markers=mm, continuity_markers=[cm] * A, refine_steps=0,
                           u_nullspace=True)
    gen = prob.solve_direct(np.zeros(space.dim()))
    next(gen)
    setup = time.perf_counter() - t0
    t1 = time.perf_counter()
    k = 0
    while k < steps and time.perf_counter() - t1 < budget_s:
        next(gen)
        k += 1
    per_step = (time.perf_counter() - t1) / max(k, 1)
    return dict(ndofs=space.dim(), per_step=per_step, setup=setup, steps=k)


def main_nullspace(args):
    """Total-pressure steps/s with a displacement nullspace (bordered FP64
    MINRES, six rigid-motion multipliers), 48^3 A = 4 traction-free."""
    import torch
    K, W = min(args.steps, 10), args.warmup
    t_setup = time.perf_counter()
    s = make_nullspace_solver(args.n, args.A)
    eng = s.engine()
    torch.cuda.synchronize()
    setup = time.perf_counter() - t_setup
    total, lat, states, its, _ = run_steps(s, W, K, keep=True)
    sysd = eng.system
    B = eng.constraints.rows()
    x = states[-1]
    cons = float((B @ x[: sysd.n]).abs().max() / x[: sysd.n].abs().max())
    xr = torch.randn(sysd.n, device="cuda", dtype=torch.float64)
    lpr = eng.krylov.csr_desc.lpr
    ms = time_kernel(lambda: sysd.csr_apply(xr, lpr=lpr, masked=False), reps=20)
    nbytes = csr_bytes(sysd)
    ach = nbytes / (ms * 1e-3) / 1e9
    out = {"metric": "MPET steps/s with a displacement nullspace (bordered FP64 MINRES, "
                     "rtol 1e-10)",
           "value": round(K / total, 4), "unit": "steps/s", "n_gpus": 1, "steps": K,
           "warmup": W, "ms_per_step": round(1e3 * total / K, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"MPET total pressure, unit cube {args.n}^3 x 6 tets, "
                                  f"A={args.A}, traction-free, u_has_nullspace, theta=1, "
                                  "dt=0.0125", "dofs": int(sysd.n),
                      "multipliers": int(eng.constraints.m)},
           "dof_per_s": round(sysd.n * K / total, 1), "iterations": its,
           "max_constraint_residual_rel": cons,
           "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                          "p99": round(1e3 * float(np.percentile(lat, 99)), 3)},
           "setup_s": round(setup, 2),
           "roofline": {"bound": "hbm", "kernel": f"csr_spmv_kernel<double,{lpr}> (A of the "
                                                  "bordered system)",
                        "achieved": round(ach, 1), "peak": 8000.0, "unit": "GB/s",
                        "frac": round(ach / 8000.0, 4), "traffic": None,
                        "bytes_per_launch": int(nbytes), "ms_per_launch": round(ms, 4)}}
    if not args.no_cpu_baseline:
        cb = nullspace_cpu_baseline(args.cpu_n, args.A, 20)
        per_full = (cb["per_step"] + cb["setup"] / 80.0) * sysd.n / cb["ndofs"]
        out["cpu_baseline"] = {
            "value": round(1.0 / per_full, 6), "unit": "steps/s (extrapolated by DOF count)",
            "cores": 1, "kind": "port",
            "sample": f"FP64 oracle bordered LU (numpy/scipy SuperLU, u nullspace) at "
                      f"{args.cpu_n}^3 ({cb['ndofs']} DOFs), {cb['steps']} steps after "
                      f"factorisation, {cb['per_step']:.3f} s/step, setup {cb['setup']:.1f} s "
                      "amortised over 80 steps"}
    print(json.dumps(out))


def main_indicators(args):
    """Two-field error indicators (AdaptiveMPETSolver, cell residuals and
    interior-facet jumps) per second on the two-field workload after two
    steps; the indicator kernels timed with HIP events."""
    import torch
    from waterscapes_amd import Expression
    from waterscapes_amd.adaptive import ErrorEstimator
    n = args.tf_n
    s = make_two_field_solver(n)
    prob = s.problem
    prob.f = Expression(("x[0]*x[1] + t", "sin(x[2])", "x[0]*x[0]"), t=prob.time, degree=2)
    f_ = Expression(("x[0]*x[1]", "sin(x[2])", "x[0]*x[0]"), degree=2)
    gen = s.solve()
    next(gen)
    next(gen)                                  # up = step 2, up_ = step 1
    est = ErrorEstimator(prob, s, f_)
    est.indicators()
    torch.cuda.synchronize()
    K = max(args.steps, 5)
    L = __import__("waterscapes_amd._lib", fromlist=["lib"])
    d = est.desc
    x, xp = s.up.data, s.up_.data

    def launch():
        L.check(L.lib().mpet_error_indicators(C.byref(d), L.ptr(x), L.ptr(xp), L.ptr(est.fac),
                                              est.nF, L.ptr(est.cell_f), L.ptr(est.tmp),
                                              L.ptr(est.out), L.stream_ptr()))
    ms = time_kernel(launch, reps=K)
    t0 = time.perf_counter()
    for _ in range(K):
        est.estimate_error_at_t()
    total = time.perf_counter() - t0
    nC = prob.mesh.num_cells()
    # algorithmic bytes: per cell the 10 P2 and 4 P1 nodes of both states,
    # geometry, the P_k source values at the cell's nodes, 4 outputs; per
    # interior facet the two cells' nodes again and 3 outputs + the gather
    A = 2
    per_cell = (2 * (30 + 4 * A) * 8 + 4 * 3 * 8 + 14 * 4 + 2 * 6 * 3 * 8 + 4 * 8)
    per_facet = (2 * (2 * (30 + 4 * A) * 8 + 4 * 3 * 8 + 14 * 4) + 4 * 4 + 3 * 8 * 2)
    nbytes = nC * per_cell + est.nF * per_facet
    ach = nbytes / (ms * 1e-3) / 1e9
    out = {"metric": "two-field error indicators per second (cells/s, GPU kernels)",
           "value": round(nC / (ms * 1e-3), 1), "unit": "cells/s", "n_gpus": 1, "steps": K,
           "warmup": 1, "ms_per_step": round(ms, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"two-field MPET, unit cube {n}^3 x 6 tets, A=2, indicators "
                                  "after two steps", "cells": nC, "interior_facets": est.nF},
           "estimate_error_at_t_ms": round(1e3 * total / K, 3),
           "roofline": {"bound": "hbm", "kernel": "ind_cell_kernel + ind_facet_kernel + "
                                                  "ind_gather_kernel",
                        "achieved": round(ach, 1), "peak": 8000.0, "unit": "GB/s",
                        "frac": round(ach / 8000.0, 4), "traffic": None,
                        "bytes_per_launch": int(nbytes), "ms_per_launch": round(ms, 4)}}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle import mpet_oracle as O
        from conftest import oracle_mesh
        from waterscapes_amd.fem import MixedSpace
        from waterscapes_amd.mesh import UnitCubeMesh
        ns = min(args.cpu_n, 6)             # the oracle's facet loop is per-facet Python
        mesh = UnitCubeMesh(ns, ns, ns)
        sp_ = MixedSpace(mesh, 2, total_pressure=False)
        m = oracle_mesh(sp_)
        rng = np.random.default_rng(1)
        xs, xps = rng.standard_normal(sp_.ndofs), rng.standard_normal(sp_.ndofs)
        fcb = lambda X, t: np.stack([X[:, 0] * X[:, 1] + t, np.sin(X[:, 2]), X[:, 0] ** 2], 1)
        t0 = time.perf_counter()
        reps = 0
        while reps < 1 or time.perf_counter() - t0 < 5.0:
            O.error_indicators(m, TF_PARAMS, xs, xps, 0.0125, f=fcb, fp=fcb, f_degree=2)
            reps += 1
            if reps >= 20:
                break
        per = (time.perf_counter() - t0) / reps
        out["cpu_baseline"] = {
            "value": round(mesh.num_cells() / per, 1), "unit": "cells/s", "cores": 1,
            "kind": "port",
            "sample": f"oracle error_indicators (numpy, FP64) at {ns}^3 ({mesh.num_cells()} "
                      f"cells), {reps} evaluations, {per:.3f} s each"}
    print(json.dumps(out))


def main_probe(args):
    """Short run for hardware-counter passes: 5 launches each of the FP32
    structured operator apply and of the cell-wise FP32 apply (cell kernel +
    row kernel, the solver's tuned lanes) of the 48^3 A = 4 workload, and 5
    launches of a plain device copy kernel of 2^27 doubles
    (copy_convert_kernel<double>: 1 GiB read, 1 GiB written) that calibrates
    the counters."""
    import torch
    from waterscapes_amd import _lib
    s = make_solver(args.n, args.A, "mixed")
    sysm = s.engine().system
    x = torch.randn(sysm.n, device="cuda", dtype=torch.float32)
    y = torch.empty_like(x)
    d32 = sysm.op_desc(1)
    L = _lib.lib()
    sp = _lib.stream_ptr()
    a = torch.randn(1 << 27, device="cuda", dtype=torch.float64)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        _lib.check(L.mpet_op_apply(C.byref(d32), _lib.ptr(x), None, _lib.ptr(y), sp))
    cd = s.engine().krylov.cell_desc
    work = torch.empty(sysm.cell_work_floats(), dtype=torch.float32, device="cuda")
    for _ in range(5):
        _lib.check(L.mpet_cell_op_apply(C.byref(cd), _lib.ptr(x), None, _lib.ptr(y),
                                        _lib.ptr(work), sp))
    for _ in range(5):
        _lib.check(L.mpet_soa_convert(0, _lib.ptr(a), 1, a.numel(), _lib.ptr(b), sp))
    torch.cuda.synchronize()
    print(json.dumps({"probe": "op_apply_fp32 x5, cell_op_apply x5, copy_convert<double> "
                               "1 GiB x5",
                      "op_bytes": int(op_bytes(sysm, 1)),
                      "cell_op_bytes": int(cell_op_bytes(sysm)),
                      "cell_lanes": [cd.lanes_u, cd.lanes_p, cd.xcd_chunk],
                      "copy_read_bytes": a.numel() * 8,
                      "copy_write_bytes": a.numel() * 8}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=48)
    ap.add_argument("--A", type=int, default=4)
    ap.add_argument("--outer-tol", type=float, default=3e-11)
    # FP32 inner MINRES tolerance: 1e-6 (two refinement steps: 191.8 ms at
    # 48^3 against 197.8 ms at 1e-5, ~200 ms at 1e-4, 201-206 ms at 1e-7 /
    # 3e-7; the 96-instance batch 62.5 instance-steps/s against 60.3 at 1e-4)
    ap.add_argument("--inner-rtol", type=float, default=None)
    # order of the time extrapolation of the initial guess, both mixed and
    # best-FP64 legs (48^3: 4 -> 12.85, 5 -> 13.14, 6 -> 13.18 steps/s over
    # the timed window; 13.65, 14.87, 15.79 over the whole 80-step run)
    ap.add_argument("--ex-order", type=int, default=6)
    # each step extrapolates with the degree (up to --ex-order) that best
    # predicted the previous step (--no-ex-adaptive: always the full order)
    ap.add_argument("--no-ex-adaptive", dest="ex_adaptive", action="store_false")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=8)
    ap.add_argument("--cpu-iters", type=int, default=4)
    # the whole configuration from t = 0 (80 steps of dt = 0.0125) on the
    # mixed path after the timed window, reported beside it (0: off)
    ap.add_argument("--full-steps", type=int, default=80)
    # the 96-instance uneven batch (20 steps each, the plain FP64 path one
    # instance at a time on a sample of 8), also reported in the default
    # line (0: off)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--batch-sample", type=int, default=8)
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-reference", action="store_true")
    ap.add_argument("--tf-n", type=int, default=32)
    ap.add_argument("--tf-precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--workload", choices=("single", "batch", "serve", "probe", "two_field",
                                           "nullspace", "indicators"),
                    default="single")
    ap.add_argument("--instances", type=int, default=96)
    ap.add_argument("--sample", type=int, default=8)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--rates", default="10,50,200,500")
    ap.add_argument("--requests", type=int, default=40)
    ap.add_argument("--steps-scale", type=float, default=1.0)
    ap.add_argument("--budget", type=float, default=200.0)
    ap.add_argument("--policy", choices=("srpt", "all"), default="srpt")
    ap.add_argument("--tick-dofs", type=float, default=6e6)
    args = ap.parse_args()
    if args.inner_rtol is None:
        args.inner_rtol = 1e-6
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: runs on one GPU only (batches would shard by instance across GPUs "
              "with no communication; one instance is not partitioned); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        return main_batch(args)
    if args.workload == "serve":
        return main_serve(args)
    if args.workload == "probe":
        return main_probe(args)
    if args.workload == "indicators":
        return main_indicators(args)
    if args.workload == "nullspace":
        return main_nullspace(args)
    if args.workload == "two_field":
        return main_two_field(args)
    import torch
    from waterscapes_amd import _lib
    K, W = args.steps, args.warmup
    t_setup = time.perf_counter()
    mixed = make_solver(args.n, args.A, "mixed", extrapolate_order=args.ex_order,
                        extrapolate_adaptive=args.ex_adaptive,
                        outer_tol=args.outer_tol,
                        inner_rtol=args.inner_rtol)
    mixed.engine()
    torch.cuda.synchronize()
    setup_mixed = time.perf_counter() - t_setup
    print(f"[bench] mixed setup {setup_mixed:.1f}s", file=sys.stderr, flush=True)
    ref_states = None
    fp64_res = None
    fp64 = None
    if not args.skip_fp64:
        t_setup = time.perf_counter()
        fp64 = make_solver(args.n, args.A, "fp64")
        fp64.engine()
        torch.cuda.synchronize()
        print(f"[bench] fp64 setup {time.perf_counter() - t_setup:.1f}s", file=sys.stderr,
              flush=True)
        fp64_res = run_steps(fp64, W, K, keep=True)
        ref_states = fp64_res[2]
        print(f"[bench] fp64 {K} steps {fp64_res[0]:.2f}s iters {fp64_res[3]}", file=sys.stderr,
              flush=True)
    best_res = None
    if not args.skip_fp64:
        # the best FP64 GPU path: the structured FP64 operator, hipGraph
        # chunks with a host check every 10 iterations, the mixed path's
        # extrapolated initial guess, and a tolerance that meets the 1e-8
        # error bar
        best = make_solver(args.n, args.A, "fp64", krylov_rtol=1e-11, extrapolate=True,
                           extrapolate_order=args.ex_order,
                           extrapolate_adaptive=args.ex_adaptive, fp64_operator="structured",
                           use_graphs=True, check_every=10)
        best.engine()
        best_res = run_steps(best, W, K, keep=True)
        print(f"[bench] fp64 best {K} steps {best_res[0]:.2f}s iters {best_res[3]}",
              file=sys.stderr, flush=True)
        best_eng = best.engine()
    mix_res = run_steps(mixed, W, K, keep=True)
    ex_orders = list(mixed.engine().ex_used[-K:]) if args.ex_adaptive else None
    print(f"[bench] mixed {K} steps {mix_res[0]:.2f}s inner {mix_res[3]} outer {mix_res[4]}",
          file=sys.stderr, flush=True)
    eng = mixed.engine()
    space = eng.space
    # live roofline of the dominant operator: the cell-wise FP32 apply of
    # the inner MINRES (cell kernel + row kernel, at the solver's tuned
    # lanes), and the FP32 structured apply it replaced; measured right
    # after the mixed leg's timed window, before the other legs
    sysm = eng.system
    x32 = (mix_res[2][-1] if mix_res[2] else mixed.up.data).float()
    y32 = torch.empty_like(x32)
    L = _lib.lib()
    sp = _lib.stream_ptr()
    cdesc = eng.krylov.cell_desc
    if cdesc is not None:
        work = torch.empty(sysm.cell_work_floats(), dtype=torch.float32, device=x32.device)
        ms = time_kernel(lambda: L.mpet_cell_op_apply(C.byref(cdesc), _lib.ptr(x32), None,
                                                      _lib.ptr(y32), _lib.ptr(work), sp))
        abytes = cell_op_bytes(sysm)
        rkernel = (f"cw_cell_kernel<256> + cw_rows_kernel<{sysm.F},{cdesc.lanes_u},"
                   f"{cdesc.lanes_p}> (cell-wise A, FP32)")
    d32 = sysm.op_desc(1)
    ms_s = time_kernel(lambda: L.mpet_op_apply(C.byref(d32), _lib.ptr(x32), None, _lib.ptr(y32),
                                               sp))
    bytes_s = op_bytes(sysm, 1)
    if cdesc is None:
        ms, abytes = ms_s, bytes_s
        rkernel = f"op_apply_kernel<float,{sysm.F},16,16> (structured A, FP32)"
    achieved = abytes / (ms * 1e-3) / 1e9
    errs, errs64, errs_best = [], [], []
    if not args.skip_reference:
        import gc
        ref = make_solver(args.n, args.A, "fp64", krylov_rtol=1e-13)
        ref_res = run_steps(ref, W, K, keep=True)
        print(f"[bench] reference rtol 1e-13 iters {ref_res[3]}", file=sys.stderr, flush=True)

        def step_err(a, b):
            e = 0.0
            for i in range(space.A + 2):
                idx = torch.as_tensor(space.field_index(i), device=a.device)
                num = torch.linalg.norm(a[idx] - b[idx]).item()
                den = torch.linalg.norm(b[idx]).item()
                e = max(e, num / den if den > 0 else num)
            return e
        errs = [step_err(a, b) for a, b in zip(mix_res[2], ref_res[2])]
        if ref_states is not None:
            errs64 = [step_err(a, b) for a, b in zip(ref_states, ref_res[2])]
        errs_best = [step_err(a, b) for a, b in zip(best_res[2], ref_res[2])] \
            if best_res is not None else []
        del ref, ref_res
        gc.collect()
    full = None
    if args.full_steps > 0:
        # the full run: zero initial state at t = 0, T = full_steps * dt
        mixed.up_.data.zero_()
        mixed.problem.time.assign(0.0)
        mixed.params["T"] = args.full_steps * 0.0125
        fres = run_steps(mixed, 0, args.full_steps)
        full = {"steps": args.full_steps, "seconds": round(fres[0], 3),
                "steps_per_s": round(args.full_steps / fres[0], 4),
                "mean_inner_iterations": round(float(np.mean(fres[3])), 1),
                "max_outer_iterations": int(max(fres[4]))}
        print(f"[bench] mixed full run {args.full_steps} steps {fres[0]:.2f}s", file=sys.stderr,
              flush=True)
    mix_sps = K / mix_res[0]
    lat = np.array(mix_res[1]) * 1e3

    def bytes_report(e, res, mixed_path):
        """Algorithmic bytes per MINRES iteration and per step of a leg, and
        the HBM fraction of its measured step time."""
        it_b, parts = iteration_bytes(e, mixed_path)
        inner = res[3]
        outer = res[4] if mixed_path else [1] * len(inner)
        st_b = [step_bytes(e, mixed_path, i, o) for i, o in zip(inner, outer)]
        tot_b = float(sum(st_b))
        secs = float(res[0])
        return {"bytes_per_iteration": it_b, "iteration_parts": parts,
                "mean_bytes_per_step": int(tot_b / len(st_b)),
                "achieved_GB_s": round(tot_b / secs / 1e9, 1),
                "frac": round(tot_b / secs / 8e12, 4)}
    out = {
        "metric": "MPET steps/s at <=1e-8 rel L2 err, mixed vs FP64 GPU; p50/p99 solve latency",
        "value": round(mix_sps, 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(1e3 * mix_res[0] / K, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-in-fp64",
        "data": "synthetic",
        "config": {"workload": f"MPET total pressure, unit cube {args.n}^3 x 6 tets, "
                               f"A={args.A} networks, theta=1, dt=0.0125",
                   "cells": 6 * args.n ** 3, "dofs": int(space.ndofs),
                   "outer_tol": args.outer_tol, "inner_rtol": args.inner_rtol},
        "dof_per_s": round(mix_sps * space.ndofs, 1),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3)},
        "inner_iterations": mix_res[3],
        "outer_iterations": mix_res[4],
        "extrapolation_orders": ex_orders,
        "max_rel_l2_err_vs_fp64_reference": max(errs) if errs else None,
        "hbm": bytes_report(eng, mix_res, True),
        "full_run": full,
        "autotune": {"picks": {str(k): v for k, v in eng.krylov.tuning.items()},
                     "legend": "operator: [lanes per P2 row, lanes per P1 row, xcd_chunk]; "
                               "levels: [A, P, R] lanes per row per level (0: unused)"},
    }
    if fp64_res is not None:
        fp_sps = K / fp64_res[0]
        out["fp64"] = {"steps_per_s": round(fp_sps, 4), "ms_per_step": round(1e3 * fp64_res[0] / K, 3),
                       "iterations": fp64_res[3], "rtol": 1e-10,
                       "operator": "scalar CSR (FP64)", "initial_guess": "up_ (as the reference)",
                       "max_rel_l2_err_vs_fp64_reference": max(errs64) if errs64 else None,
                       "hbm": bytes_report(fp64.engine(), fp64_res, False)}
        out["speedup_mixed_vs_fp64"] = round(mix_sps / fp_sps, 3)
        if best_res is not None:
            fb_sps = K / best_res[0]
            out["fp64_best"] = {
                "steps_per_s": round(fb_sps, 4), "ms_per_step": round(1e3 * best_res[0] / K, 3),
                "iterations": best_res[3], "rtol": 1e-11,
                "operator": "structured node-pattern operator (FP64), hipGraph chunks, "
                            "host check every 10 iterations",
                "initial_guess": f"extrapolation of order {args.ex_order}"
                                 f"{' (adaptive)' if args.ex_adaptive else ''} (as the mixed path)",
                "max_rel_l2_err_vs_fp64_reference": max(errs_best) if errs_best else None,
                "hbm": bytes_report(best_eng, best_res, False)}
            out["speedup_mixed_vs_fp64_best"] = round(mix_sps / fb_sps, 3)
        fsys = fp64.engine().system
        if fsys.csr is not None:
            rp, ci, v = fsys.csr
            x64 = ref_states[-1]
            y64 = torch.empty_like(x64)
            lpr = fp64.engine().krylov.csr_desc.lpr
            ms64 = time_kernel(lambda: L.mpet_csr_spmv(0, fsys.n, _lib.ptr(rp), _lib.ptr(ci),
                                                       _lib.ptr(v), _lib.ptr(x64), None, None,
                                                       _lib.ptr(y64), lpr, sp))
            out["fp64"]["csr_spmv"] = {"ms": round(ms64, 4), "GB_s": round(csr_bytes(fsys) / (ms64 * 1e-3) / 1e9, 1)}
    def traffic_of(name):
        # HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE counter passes
        # (calibrated on a copy kernel of known size), recorded in profiles/
        tfile = os.path.join(ROOT, "profiles", name)
        if args.n == 48 and args.A == 4 and os.path.exists(tfile):
            return int(json.load(open(tfile))["traffic_bytes_per_launch"])
        return None
    out["roofline"] = {"bound": "hbm", "kernel": rkernel,
                       "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                       "frac": round(achieved / 8000.0, 4),
                       "traffic": traffic_of("cell_op_48cube_A4_traffic.json" if cdesc is not None
                                             else "op_apply_48cube_A4_traffic.json"),
                       "bytes_per_launch": int(abytes), "ms_per_launch": round(ms, 4)}
    if cdesc is not None:
        ach_s = bytes_s / (ms_s * 1e-3) / 1e9
        out["roofline_structured_operator"] = {
            "kernel": f"op_apply_kernel<float,{sysm.F},16,16> (assembled A, FP32; replaced in "
                      f"the inner MINRES by the cell-wise apply)",
            "achieved": round(ach_s, 1), "frac": round(ach_s / 8000.0, 4),
            "traffic": traffic_of("op_apply_48cube_A4_traffic.json"),
            "bytes_per_launch": int(bytes_s), "ms_per_launch": round(ms_s, 4)}
    # the finest displacement level's Chebyshev step (the top kernel by time
    # in the V-cycles): half values and 16-bit offsets, three columns, the
    # solver's tuned lanes; bytes as vcycle_bytes counts a middle step
    # (direction read and written)
    kr = eng.krylov
    if getattr(kr, "compact", None) is not None:
        from waterscapes_amd import autotune as at
        (ulevs, _), _p = kr._level_arrays(1)
        ld = ulevs[0]
        n0 = int(kr.levels[0]["n"])
        nnz0 = int(ld["A_ci"].numel())
        lpr0 = int(kr.tuning.get(1, {}).get("u_levels", [[4]])[0][0] or 4)
        xs = torch.randn(3 * n0, device=x32.device, dtype=torch.float32)
        bs = torch.randn_like(xs)
        ys = torch.empty_like(xs)
        ds = torch.zeros_like(xs)
        Lt = at._bind()
        ms_v = time_kernel(lambda: Lt.mpet_level_spmm(
            1, 3, 3, n0, _lib.ptr(ld["A_rp"]), _lib.ptr(ld["A_ci"]), None, _lib.ptr(xs),
            _lib.ptr(bs), _lib.ptr(ys), _lib.ptr(ld["dinv"]), _lib.ptr(ds), 0.5, 0.5,
            _lib.ptr(kr.mask0), lpr0, _lib.ptr(ld["A_v16"]), _lib.ptr(ld["A_dci"]),
            float(ld["vscale"]), sp))
        bytes_v = nnz0 * 4 + (n0 + 1) * 4 + 3 * n0 * 4 + 4 * 3 * n0 * 4 + n0 * 4
        ach_v = bytes_v / (ms_v * 1e-3) / 1e9
        out["roofline_vcycle_level0"] = {
            "kernel": f"spmm_kernel<float,3,{lpr0},3,true> (P2-level Chebyshev step of the "
                      f"displacement V-cycle, half values, {nnz0} nonzeros x 3 columns)",
            "achieved": round(ach_v, 1), "frac": round(ach_v / 8000.0, 4),
            "bytes_per_launch": int(bytes_v), "ms_per_launch": round(ms_v, 4)}
    if not args.no_cpu_baseline and fp64_res is not None:
        # the plain FP64 path on the host at the full workload, a bounded
        # sample of MINRES iterations scaled by the iterations per step
        its = float(np.mean(fp64_res[3]))
        cb = cpu_baseline_headline(fp64.engine(), its, iters=args.cpu_iters)
        out["cpu_baseline"] = {
            "value": round(1.0 / cb["per_step"], 6), "unit": "steps/s", "cores": 1,
            "kind": "port",
            "sample": f"FP64 MINRES of the full {args.n}^3 workload on the host (the oracle's "
                      f"textbook MINRES, the same block preconditioner restated in numpy on the "
                      f"same GPU-assembled CSR matrix and host AMG hierarchies, scipy sparse "
                      f"products): {cb['iters']} iterations of one step plus its start "
                      f"in {cb['wall']:.1f} s ({cb['per_iteration']:.2f} s per iteration) x "
                      f"{its:.1f} iterations per step (the GPU FP64 path's mean at the same "
                      f"rtol 1e-10); single-threaded sparse products, "
                      f"{len(os.sched_getaffinity(0))} threads visible"}
    if args.batch_steps > 0 and not args.skip_fp64:
        bt = run_batch(args, 96, args.batch_steps, 2, args.batch_sample)
        out["batch"] = {k: bt[k] for k in ("metric", "value", "unit", "steps", "ms_per_step",
                                           "dof_steps_per_s", "setup_s", "fp64_one_at_a_time",
                                           "speedup_vs_fp64_one_at_a_time")}
        out["batch"]["config"] = bt["config"]["workload"]
        print(f"[bench] batch of 96: {bt['value']} instance-steps/s, "
              f"{bt['speedup_vs_fp64_one_at_a_time']}x", file=sys.stderr, flush=True)
    print(json.dumps(out))

if __name__ == "__main__":
    main()
