"""Device Krylov solver of one MPET instance: MINRES + block preconditioner
(csrc/krylov.hip), FP64 or FP32-in-FP64-refinement.

Host work done here once per instance (setup): Dirichlet masks, the
displacement-block hierarchy (native SA setup on the eliminated P2
Laplacian, first coarsening by the exact P1 -> P2 embedding), spectral bounds
of the pressure blocks, and the upload of every array to HBM.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib, amg, forms
from .amg import AmgParams
from ._lib import check, ptr

MAXLEV = 16


class AmgLevelDesc(C.Structure):
    _fields_ = [("n", C.c_int32), ("nc", C.c_int32), ("A_rp", C.c_void_p), ("A_ci", C.c_void_p),
                ("A_v", C.c_void_p), ("dinv", C.c_void_p), ("rho", C.c_double),
                ("P_rp", C.c_void_p), ("P_ci", C.c_void_p), ("P_v", C.c_void_p),
                ("R_rp", C.c_void_p), ("R_ci", C.c_void_p), ("R_v", C.c_void_p),
                ("lpr", C.c_int32), ("lprP", C.c_int32), ("lprR", C.c_int32),
                ("A_v16", C.c_void_p), ("A_dci", C.c_void_p), ("A_vscale", C.c_double),
                ("grp", C.c_void_p)]


class HierDesc(C.Structure):
    _fields_ = [("nlev", C.c_int32), ("lev", AmgLevelDesc * MAXLEV), ("coarse_inv", C.c_void_p),
                ("degree", C.c_int32), ("eig_ratio", C.c_double), ("eig_safety", C.c_double),
                ("scale", C.c_double), ("mask0", C.c_void_p), ("coarse_blk", C.c_void_p),
                ("coarse_off", C.c_void_p), ("coarse_ioff", C.c_void_p),
                ("level_degree", C.c_int32 * MAXLEV)]


class PrecDesc(C.Structure):
    _fields_ = [("u", HierDesc), ("p", HierDesc), ("nP1", C.c_int32), ("F", C.c_int32),
                ("pmap", C.c_void_p)]


class CsrDesc(C.Structure):
    _fields_ = [("nrows", C.c_longlong), ("rp", C.c_void_p), ("ci", C.c_void_p),
                ("v", C.c_void_p), ("lpr", C.c_int32)]


class SolverOptions(C.Structure):
    _fields_ = [("mode", C.c_int32), ("use_csr", C.c_int32), ("rtol", C.c_double),
                ("maxit", C.c_int32), ("check_every", C.c_int32), ("use_graphs", C.c_int32),
                ("outer_tol", C.c_double), ("inner_rtol", C.c_double), ("max_outer", C.c_int32)]


class SolveStats(C.Structure):
    _fields_ = [("inner_iterations", C.c_int32), ("outer_iterations", C.c_int32),
                ("converged", C.c_int32), ("residual_norm", C.c_double)]


def _bind():
    L = _lib.lib()
    P = C.c_void_p
    L.mpet_solver_create.argtypes = [P, P, P, P, P, P, C.POINTER(C.c_void_p)]
    L.mpet_solver_create.restype = C.c_int
    L.mpet_solver_create_csr.argtypes = [P, P, C.c_int, C.c_int, C.c_int, P, P,
                                         C.POINTER(C.c_void_p)]
    L.mpet_solver_create_csr.restype = C.c_int
    L.mpet_solver_create_bordered.argtypes = [P, P, C.c_int, C.c_int, C.c_int, P, P, C.c_int,
                                              C.c_double, P, C.POINTER(C.c_void_p)]
    L.mpet_solver_create_bordered.restype = C.c_int
    L.mpet_solver_create_csr_mixed.argtypes = [P, P, C.c_int, C.c_int, C.c_int, P, P,
                                               C.POINTER(C.c_void_p)]
    L.mpet_solver_create_csr_mixed.restype = C.c_int
    L.mpet_solver_destroy.argtypes = [P]
    L.mpet_solver_destroy.restype = C.c_int
    L.mpet_solver_solve.argtypes = [P, P, P, P, P]
    L.mpet_solver_solve.restype = C.c_int
    L.mpet_solver_set_cell_operator.argtypes = [P, P]
    L.mpet_solver_set_cell_operator.restype = C.c_int
    L.mpet_solver_set_cell_operator64.argtypes = [P, P]
    L.mpet_solver_set_cell_operator64.restype = C.c_int
    L.mpet_precond_apply.argtypes = [P, C.c_int, P, P, P]
    L.mpet_precond_apply.restype = C.c_int
    return L


@dataclass
class SolverConfig:
    """Solver settings.  mode 'fp64': plain FP64 MINRES to rtol on the
    preconditioned residual (host check every iteration, like PETSc KSP);
    mode 'mixed': FP32 MINRES corrections inside FP64 refinement to
    outer_tol relative FP64 residual, checked every check_every iterations
    from captured hipGraph chunks."""
    mode: str = "fp64"
    rtol: float = 1e-10
    maxit: int = 5000
    outer_tol: float = 3e-11
    inner_rtol: float = 1e-6
    max_outer: int = 30
    check_every: int = 0              # 0: 1 for fp64, 10 for mixed
    use_graphs: bool = None           # default: mixed only
    use_csr: bool = None              # default: fp64 only
    # displacement hierarchy: one Chebyshev step on the P2 level (as many
    # MINRES iterations as two at 48^3, measured), three on the coarser ones
    amg: AmgParams = field(default_factory=lambda: AmgParams(smoother_degree=3,
                                                             level_degrees=(1,)))
    pamg: AmgParams = field(default_factory=AmgParams)
    # weight of the total-pressure mass in the preconditioner.  The reference
    # form prec uses the bare mass p0 w0 dx (mpettotalpressuresolver.py:282),
    # which is off from the p0 Schur complement ~ (1/(2 mu) + 1/lambda) M by
    # a factor ~mu and costs ~4x MINRES iterations at E = 1500; "robust"
    # uses 1/(2 mu) + 1/lambda, "reference" the bare mass.
    p0_weight: str = "robust"
    # FP32 hierarchies: keep the finest displacement-level operator as half
    # precision values (scaled into half's range) with 16-bit column offsets,
    # half the bytes of its smoothing / residual SpMMs; the V-cycle stays a
    # fixed symmetric operator, so MINRES is unaffected beyond the slightly
    # perturbed preconditioner
    compact_level0: bool = True
    # the same for the pressure levels (the finest: field-major
    # block-diagonal P_f) but the dense coarsest, each when every row's
    # diagonal stays in half's normal range under one scale (the coarser
    # ones only with the finest)
    compact_plevel0: bool = True
    # launch shapes picked by the autotuner (waterscapes_amd/autotune.py):
    # None = on for systems of at least autotune_min_dofs unknowns (below,
    # launches are latency-bound and the fixed heuristic is kept)
    autotune: bool = None
    autotune_min_dofs: int = 200000
    # inner FP32 MINRES of the mixed mode on the cell-wise (matrix-free)
    # operator instead of the FP32 copy of the assembled one
    cell_operator: bool = True
    # the mixed mode's FP64 residuals (refinement) through the cell-wise
    # FP64 operator (same formulas, FP64 geometry) instead of the assembled
    # FP64 values: 2.6 GB fewer bytes per residual at 48^3
    cell_operator64: bool = True


def set_level_degrees(h, params: AmgParams):
    """Per-level Chebyshev degrees of a HierDesc from the AMG parameters."""
    if len(params.level_degrees) > MAXLEV:
        raise ValueError("more level degrees than hierarchy levels")
    for l in range(MAXLEV):
        h.level_degree[l] = params.degree(l)


def _csr_i32(rp):
    if rp[-1] >= 2**31:
        raise ValueError("level too large for int32 row pointers")
    return rp.astype(np.int32)


def _pick_lpr(avg, nc=1, rows=None):
    """Lanes per row: about four nonzeros per lane on large scalar levels and
    eight on three-column levels (fewer cross-lane reductions, more loads in
    flight per lane; measured fastest on the 48^3 levels), more lanes while
    a small level would leave the chip under-filled (fewer than ~1M lanes in
    flight, ~0.5M for three columns); at least one lane per component."""
    per_lane, fill = (8.0, 1 << 19) if nc == 3 else (4.0, 1 << 20)
    target = max(avg / per_lane, 1.0)
    lanes = 1
    while lanes < target and lanes < 64:
        lanes *= 2
    if rows is not None:
        while rows * lanes < fill and lanes < 64 and lanes < avg:
            lanes *= 2
    if nc == 3 and lanes < 4:
        # three-column rows of a few nonzeros (the prolongator from the P1
        # level): one thread per row; otherwise one lane per component
        return 1 if avg <= 4.0 else 4
    return lanes


def compact_level(pattern_host, values, min_diag_ratio=None):
    """Half-precision copy of a level operator on a host CSR pattern
    (rp, ci) with device FP64 values: values * 2^e rounded to half (the
    largest magnitude mapped near 2^10), columns as int16 offsets from the
    row.  None if some offset does not fit in 16 bits, or (min_diag_ratio
    given) if some nonzero diagonal entry is smaller than min_diag_ratio
    times the largest value: half is absolutely accurate to 2^-25 (scaled)
    below its normal range, so with every nonzero diagonal normal each
    row's entries keep half's relative precision measured on the row."""
    rp, ci = pattern_host
    rows = np.repeat(np.arange(rp.size - 1, dtype=np.int64), np.diff(rp))
    off = ci.astype(np.int64) - rows
    if off.size == 0 or np.abs(off).max() > 32767:
        return None
    mag = values.abs()
    vmax = float(mag.max().item())
    if not vmax > 0:
        return None
    if min_diag_ratio is not None:
        dg = mag[torch.as_tensor(np.flatnonzero(off == 0)).to(mag.device)]
        dg = dg[dg > 0]
        if dg.numel() == 0 or float(dg.min().item()) < min_diag_ratio * vmax:
            return None
    e = 10 - int(np.ceil(np.log2(vmax)))
    v16 = (values * (2.0 ** e)).to(torch.float16)
    dci = torch.as_tensor(off.astype(np.int16)).to(values.device)
    return {"A_v16": v16, "A_dci": dci, "vscale": 2.0 ** (-e), "exp": e}


class KrylovSolver:
    """Owns the device preconditioner data and the native solver handle."""

    def __init__(self, system, bc_dofs, config: SolverConfig = None, create_handle=True,
                 constraints=None):
        self.sys = system
        self.cfg = cfg = config or SolverConfig()
        # nullspace multipliers (nullspace.Constraints): bordered FP64 system
        self.cons = constraints
        if constraints is not None and cfg.mode != "fp64":
            raise ValueError("systems with nullspace multipliers are solved in FP64 "
                             "(mode 'fp64') only")
        dev = system.device
        self.keep = []                       # tensors the native handle borrows
        if cfg.mode not in ("fp64", "mixed", "fp64-pc32"):
            raise ValueError("mode must be 'fp64', 'mixed' or 'fp64-pc32'")
        mixed = cfg.mode == "mixed"
        pc32 = cfg.mode in ("mixed", "fp64-pc32")
        # Dirichlet masks
        dofs = np.unique(np.concatenate([np.asarray(d, dtype=np.int64) for d in bc_dofs])) \
            if len(bc_dofs) else np.zeros(0, np.int64)
        system.set_dirichlet(dofs)
        nu = system.nu
        udofs = dofs[dofs < nu]
        nodes = np.unique(udofs // 3)
        for j in range(3):
            if not np.array_equal(np.unique(udofs[udofs % 3 == j] // 3), nodes):
                raise ValueError("Dirichlet conditions on u must constrain all components")
        self.dir_nodes = nodes
        mask0 = np.zeros(system.nP2, dtype=np.uint8)
        mask0[nodes] = 1
        self.mask0 = torch.as_tensor(mask0).to(dev)
        pmask = np.zeros(system.F * system.nP1, dtype=np.uint8)
        pd = dofs[dofs >= nu] - nu
        pmask[pd] = 1
        self.pmask = torch.as_tensor(pmask).to(dev)
        self.Pd = np.array(system.Pd, dtype=np.float64)
        two_field = getattr(system, "two_field", False)
        if two_field and cfg.mode not in ("fp64", "mixed"):
            raise ValueError("the two-field system is solved in mode 'fp64' or 'mixed'")
        if cfg.p0_weight not in ("robust", "reference"):
            raise ValueError("p0_weight must be 'robust' or 'reference'")
        if cfg.p0_weight == "robust" and not two_field:
            self.Pd[0] = 0.5 / system.mu + 1.0 / system.lmbda
        if constraints is not None:
            # constrained pressure fields: the reference adds p[k] q[k] dx
            # (mpettotalpressuresolver.py:304-305, mpetsolver.py:274) so that
            # a pure-Neumann block without a mass term stays definite; a block
            # that already has one keeps its weight (adding the bare mass to
            # the "robust" p_0 weight ~1/(2 mu) would swamp it)
            for f in constraints.fields:
                if not self.Pd[f] > 0.0:
                    self.Pd[f] = 1.0
        t0 = time.perf_counter()
        self._build_hierarchy()
        t1 = time.perf_counter()
        self._build_pressure()
        self.setup_seconds = {"u_hierarchy": t1 - t0, "p_hierarchy": time.perf_counter() - t1}
        self.handle = None
        self.last_stats = None
        if not create_handle:        # host hierarchies only (packed by a batch)
            return
        L = _bind()
        self.prec64 = self._prec_desc(0) if not pc32 else None
        self.prec32 = self._prec_desc(1) if pc32 else None
        use_csr = cfg.use_csr if cfg.use_csr is not None else (cfg.mode == "fp64")
        if two_field or constraints is not None:
            use_csr = True
        self.csr_desc = None
        if use_csr and (cfg.mode == "fp64" or two_field):
            if system.csr is None:
                system.csr = system.expand_csr(0)
            rp, ci, v = system.csr
            avg = (system.nnz_csr if hasattr(system, "nnz_csr") else v.numel()) / system.n
            # the two-field rows (~100 nonzeros per u row, ~215 per p row) run
            # fastest at 16 lanes (measured at 32^3: 0.277 vs 0.312 ms at 32)
            lpr = 16 if two_field else min(32, _pick_lpr(avg))
            self.csr_desc = CsrDesc(system.n, ptr(rp), ptr(ci), ptr(v), lpr)
        check_every = cfg.check_every or (1 if cfg.mode == "fp64" else 10)
        graphs = cfg.use_graphs if cfg.use_graphs is not None else (cfg.mode != "fp64")
        mode_id = {"fp64": 0, "mixed": 1, "fp64-pc32": 2}[cfg.mode]
        self.opt = SolverOptions(mode_id, 1 if self.csr_desc is not None else 0,
                                 cfg.rtol, cfg.maxit, check_every, 1 if graphs else 0,
                                 cfg.outer_tol, cfg.inner_rtol, cfg.max_outer)
        h = C.c_void_p()
        if constraints is not None:
            self.op64 = self.op32 = None
            check(L.mpet_solver_create_bordered(C.byref(self.csr_desc), ptr(system.dmask),
                                                system.nP2, system.nP1, system.F,
                                                C.byref(self.prec64), ptr(constraints.B),
                                                constraints.m, constraints.omega,
                                                C.byref(self.opt), C.byref(h)))
            self.handle = h
            return
        if two_field:
            self.op64 = self.op32 = None
            if mixed:
                check(L.mpet_solver_create_csr_mixed(C.byref(self.csr_desc), ptr(system.dmask),
                                                     system.nP2, system.nP1, system.F,
                                                     C.byref(self.prec32), C.byref(self.opt),
                                                     C.byref(h)))
            else:
                check(L.mpet_solver_create_csr(C.byref(self.csr_desc), ptr(system.dmask),
                                               system.nP2, system.nP1, system.F,
                                               C.byref(self.prec64), C.byref(self.opt),
                                               C.byref(h)))
            self.handle = h
            return
        self.op64 = system.op_desc(0)
        self.op32 = system.op_desc(1) if mixed else None
        check(L.mpet_solver_create(C.byref(self.op64),
                                   C.byref(self.op32) if self.op32 is not None else None,
                                   C.byref(self.csr_desc) if self.csr_desc is not None else None,
                                   C.byref(self.prec64) if self.prec64 is not None else None,
                                   C.byref(self.prec32) if self.prec32 is not None else None,
                                   C.byref(self.opt), C.byref(h)))
        self.handle = h
        self.last_stats = None
        self.cell_desc = None
        self.cell_desc64 = None
        if mixed and cfg.cell_operator:
            self.cell_desc = system.cell_op_desc()
            check(L.mpet_solver_set_cell_operator(h, C.byref(self.cell_desc)))
            if cfg.cell_operator64:
                self.cell_desc64 = system.cell_op_desc(fp64=True)
                check(L.mpet_solver_set_cell_operator64(h, C.byref(self.cell_desc64)))
        self.tuning = {}
        tune = cfg.autotune if cfg.autotune is not None else system.n >= cfg.autotune_min_dofs
        if tune:
            self.autotune()

    def level_heuristic(self, levels, levs, nc):
        """The fixed lanes-per-row rule (_pick_lpr) of level l's operator,
        prolongator or restriction: the autotuner's incumbent."""
        def rule(l, which):
            lev, ld = levels[l], levs[l]
            if which == "A":
                return _pick_lpr(ld["A_ci"].numel() / max(lev["n"], 1), nc, lev["n"])
            nnzp = ld["P_ci"].numel()
            if which == "P":
                return _pick_lpr(nnzp / max(lev["n"], 1), nc, lev["n"])
            return _pick_lpr(nnzp / max(lev["nc"], 1), nc, lev["nc"])
        return rule

    def autotune(self, cache=None):
        """Pick and apply the launch shapes of the operator(s) and level
        SpMMs this solver runs (waterscapes_amd/autotune.py); the picks are
        kept in self.tuning."""
        from . import autotune as at
        s = self.sys
        dtypes = []
        if self.prec32 is not None:
            dtypes.append(1)
        if self.prec64 is not None:
            dtypes.append(0)
        for dt in dtypes:
            op = None
            # the FP32 structured operator runs only when the cell-wise one is off
            structured = (self.csr_desc is None or dt == 1) and not (
                dt == 1 and self.cell_desc is not None)
            if structured:
                op = at.tune_operator(s, dt, cache)
            (ulevs, _), (plevs, _) = self._level_arrays(dt)
            u = at.tune_hierarchy(self.levels, ulevs, 3, dt,
                                  self.level_heuristic(self.levels, ulevs, 3), self.mask0, cache)
            p = at.tune_hierarchy(self.plevels, plevs, 1, dt,
                                  self.level_heuristic(self.plevels, plevs, 1), None, cache)
            at.apply_tuning(self.handle, dt, op=op, u=u, p=p)
            self.tuning[dt] = {"operator": list(op) if op else None, "u_levels": u,
                               "p_levels": p}
        if self.cell_desc is not None:
            cell = at.tune_cell_operator(s, cache)
            self.cell_desc = s.cell_op_desc(lanes=(cell[0], cell[1]), xcd_chunk=cell[2])
            check(_bind().mpet_solver_set_cell_operator(self.handle, C.byref(self.cell_desc)))
            self.tuning["cell"] = list(cell)
            if self.cell_desc64 is not None:
                # the FP64 residuals: the same tables and launch shapes
                self.cell_desc64 = s.cell_op_desc(lanes=(cell[0], cell[1]), xcd_chunk=cell[2],
                                                  fp64=True)
                check(_bind().mpet_solver_set_cell_operator64(self.handle,
                                                              C.byref(self.cell_desc64)))
        if self.prec32 is not None and 1 in self.tuning and self.opt.mode in (1, 2) and \
                self.cell_desc64 is None:
            # the FP64 residual of the refinement runs the FP64 operator
            op64 = at.tune_operator(s, 0, cache)
            at.apply_tuning(self.handle, 0, op=op64)
            self.tuning["op64"] = list(op64)
        return self.tuning

    # -- setup ----------------------------------------------------------------
    def _build_hierarchy(self):
        s = self.sys
        rp, ci = s.pat_host["22"]
        lap_t = s.prim_t["Lap"]
        if self.cons is not None and self.cons.M2 is not None:
            # displacement nullspace: prec += inner(u, v) dx
            # (mpettotalpressuresolver.py:290-293); the hierarchy is scaled by mu
            lap_t = lap_t + self.cons.M2 / s.mu
        self.lap0 = lap_t
        lap = lap_t.cpu().numpy()
        L = sp.csr_matrix((lap, ci, rp), shape=(s.nP2, s.nP2))
        m = np.zeros(s.nP2, dtype=bool)
        m[self.dir_nodes] = True
        keep = sp.diags((~m).astype(np.float64))
        Le = (keep @ L @ keep + sp.diags(m.astype(np.float64))).tocsr()
        Le.sort_indices()
        P0 = amg.p1_embedding(s.space, self.dir_nodes)
        self.levels, self.cinv = amg.setup_host(Le.indptr, Le.indices, Le.data,
                                                self.cfg.amg, P0=P0)
        # level 0 on the device is the un-eliminated Lap (output masked)
        self.level_dev = {}

    def _upload_levels(self, levels, cinv, dtype, level0=None):
        """Device copies of a host hierarchy; level0 = (rp, ci, v) device
        tensors to use instead of the host level-0 matrix."""
        dev = self.sys.device
        tdt = torch.float64 if dtype == 0 else torch.float32
        out = []
        for l, lev in enumerate(levels):
            d = {}
            if l == 0 and level0 is not None:
                d["A_rp"], d["A_ci"], d["A_v"] = level0
            else:
                d["A_rp"] = torch.as_tensor(_csr_i32(lev["A_rp"])).to(dev)
                d["A_ci"] = torch.as_tensor(lev["A_ci"]).to(dev)
                d["A_v"] = torch.as_tensor(lev["A_v"]).to(dev, tdt)
            d["dinv"] = torch.as_tensor(lev["dinv"]).to(dev, tdt)
            if lev["nc"]:
                for k in ("P", "R"):
                    d[k + "_rp"] = torch.as_tensor(_csr_i32(lev[k + "_rp"])).to(dev)
                    d[k + "_ci"] = torch.as_tensor(lev[k + "_ci"]).to(dev)
                    d[k + "_v"] = torch.as_tensor(lev[k + "_v"]).to(dev, tdt)
            out.append(d)
        return out, torch.as_tensor(cinv).to(dev, tdt)

    def _level_arrays(self, dtype):
        if dtype in self.level_dev:
            return self.level_dev[dtype]
        s = self.sys
        lap = self.lap0 if dtype == 0 else self.lap0.to(torch.float32)
        u = self._upload_levels(self.levels, self.cinv, dtype,
                                level0=(s.pat_dev["22"][0], s.pat_dev["22"][1], lap))
        if dtype == 1 and self.cfg.compact_level0:
            c = compact_level(s.pat_host["22"], self.lap0)
            if c is not None:
                u[0][0].update(c)
                self.compact = c
        p = self._upload_levels(self.plevels, self.pcinv, dtype)
        if dtype == 1 and self.cfg.compact_plevel0:
            # half keeps 11 bits down to 2^-14: with the largest value near
            # 2^10, diagonals down to 2^-22 of it stay normal
            # (Dirichlet rows, identity rows of the eliminated P_f, are
            # stored as zeros and masked: the pressure residual is zero there)
            lev = self.plevels[0]
            n1, F = self.sys.nP1, self.sys.F
            dm = self.pmask.view(n1, F).t().contiguous().view(-1)
            rows = np.repeat(np.arange(lev["n"], dtype=np.int64), np.diff(lev["A_rp"]))
            v = torch.as_tensor(lev["A_v"]).to(self.sys.device)
            v[dm[torch.as_tensor(rows).to(v.device)].bool()] = 0.0
            c = compact_level((lev["A_rp"], lev["A_ci"]), v, 2.0 ** -22)
            self.pcompact_levels = []
            if c is not None:
                p[0][0].update(c, mask0=dm)
                self.pcompact_levels.append(0)
                # the coarser levels but the dense coarsest: their SpMMs
                # read as many bytes per V-cycle as the finest level's
                for l in range(1, len(self.plevels) - 1):
                    lv = self.plevels[l]
                    cl = compact_level((lv["A_rp"], lv["A_ci"]),
                                       torch.as_tensor(lv["A_v"]).to(self.sys.device),
                                       2.0 ** -22)
                    if cl is not None:
                        p[0][l].update(cl)
                        self.pcompact_levels.append(l)
        self.level_dev[dtype] = (u, p)
        return u, p

    def _build_pressure(self):
        """Hierarchy of the block-diagonal pressure preconditioner
        diag(P_0, .., P_A) (two-field: diag(P_1, .., P_A)), P_f = Pd_f M1 +
        Ps_f St_f (Dirichlet rows and columns eliminated), in field-major
        order."""
        s = self.sys
        rp, ci = s.pat_host["11"]
        M1 = s.prim_t["M1"].cpu().numpy()
        St = s.prim_t["StK"].cpu().numpy().reshape(s.A, -1)     # prec has no Robin term
        n1 = s.nP1
        F = s.F
        Mm = sp.csr_matrix((M1, ci, rp), shape=(n1, n1))
        pmask = self.pmask.cpu().numpy().reshape(n1, F).astype(bool)
        blocks = []
        for f in range(F):
            Pf = self.Pd[f] * Mm
            k = s.st_row(f)
            if k is not None:
                Pf = Pf + s.Ps[f] * sp.csr_matrix((St[k], ci, rp), shape=(n1, n1))
            free = ~pmask[:, f]
            keep = sp.diags(free.astype(np.float64))
            blocks.append((keep @ Pf @ keep + sp.diags((~free).astype(np.float64))).tocsr())
        Pp = sp.block_diag(blocks, format="csr")
        Pp.sort_indices()
        self.Pp_host = Pp
        self.plevels, self.pcinv = amg.setup_host(Pp.indptr, Pp.indices, Pp.data, self.cfg.pamg)

    def _hier_desc(self, h, levels, levs, cinv, params, scale, mask0, nc):
        h.nlev = len(levels)
        if h.nlev > MAXLEV:
            raise ValueError("hierarchy too deep")
        for l, (lev, ld) in enumerate(zip(levels, levs)):
            L = h.lev[l]
            L.n = lev["n"]
            L.nc = lev["nc"]
            L.A_rp, L.A_ci, L.A_v = ptr(ld["A_rp"]), ptr(ld["A_ci"]), ptr(ld["A_v"])
            L.dinv = ptr(ld["dinv"])
            L.rho = lev["rho"]
            L.lpr = _pick_lpr(ld["A_ci"].numel() / max(lev["n"], 1), nc, lev["n"])
            if "A_v16" in ld:
                L.A_v16, L.A_dci, L.A_vscale = ptr(ld["A_v16"]), ptr(ld["A_dci"]), ld["vscale"]
            if lev["nc"]:
                L.P_rp, L.P_ci, L.P_v = ptr(ld["P_rp"]), ptr(ld["P_ci"]), ptr(ld["P_v"])
                L.R_rp, L.R_ci, L.R_v = ptr(ld["R_rp"]), ptr(ld["R_ci"]), ptr(ld["R_v"])
                nnzp = ld["P_ci"].numel()
                L.lprP = _pick_lpr(nnzp / max(lev["n"], 1), nc, lev["n"])
                L.lprR = _pick_lpr(nnzp / max(lev["nc"], 1), nc, lev["nc"])
        h.coarse_inv = ptr(cinv)
        h.degree = params.smoother_degree
        set_level_degrees(h, params)
        h.eig_ratio = params.eig_ratio
        h.eig_safety = params.eig_safety
        h.scale = scale
        h.mask0 = ptr(mask0) if mask0 is not None else None

    def _prec_desc(self, dtype):
        s = self.sys
        (ulevs, ucinv), (plevs, pcinv) = self._level_arrays(dtype)
        self.keep += [ulevs, ucinv, plevs, pcinv]
        d = PrecDesc()
        self._hier_desc(d.u, self.levels, ulevs, ucinv, self.cfg.amg, s.mu, self.mask0, 3)
        self._hier_desc(d.p, self.plevels, plevs, pcinv, self.cfg.pamg, 1.0,
                        plevs[0].get("mask0"), 1)
        d.nP1 = s.nP1
        d.F = s.F
        return d

    # -- use ------------------------------------------------------------------
    def solve(self, b, x, raise_on_fail=True):
        """Solve one step: b full right-hand side (FP64 device), x initial
        guess with Dirichlet values in place (overwritten by the solution)."""
        L = _bind()
        st = SolveStats()
        rc = L.mpet_solver_solve(self.handle, ptr(b), ptr(x), C.byref(st), _lib.stream_ptr())
        self.last_stats = st
        if rc != 0 and (raise_on_fail or rc != 3):
            check(rc)
        return st

    def precond(self, r, dtype=0):
        z = torch.empty_like(r)
        check(_bind().mpet_precond_apply(self.handle, dtype, ptr(r), ptr(z), _lib.stream_ptr()))
        return z

    def close(self):
        if getattr(self, "handle", None):
            torch.cuda.synchronize()
            _bind().mpet_solver_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
