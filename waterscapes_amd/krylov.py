"""Device Krylov solver of one MPET instance: MINRES + block preconditioner
(csrc/krylov.hip), FP64 or FP32-in-FP64-refinement.

Host work done here once per instance (setup): Dirichlet masks, the
displacement-block hierarchy (native SA setup on the eliminated P2
Laplacian, first coarsening by the exact P1 -> P2 embedding), spectral bounds
of the pressure blocks, and the upload of every array to HBM.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib, amg, forms
from ._lib import check, ptr

MAXLEV = 16


class AmgLevelDesc(C.Structure):
    _fields_ = [("n", C.c_int32), ("nc", C.c_int32), ("A_rp", C.c_void_p), ("A_ci", C.c_void_p),
                ("A_v", C.c_void_p), ("dinv", C.c_void_p), ("rho", C.c_double),
                ("P_rp", C.c_void_p), ("P_ci", C.c_void_p), ("P_v", C.c_void_p),
                ("R_rp", C.c_void_p), ("R_ci", C.c_void_p), ("R_v", C.c_void_p),
                ("lpr", C.c_int32), ("lprP", C.c_int32), ("lprR", C.c_int32)]


class PrecDesc(C.Structure):
    _fields_ = [("nlev", C.c_int32), ("lev", AmgLevelDesc * MAXLEV), ("coarse_inv", C.c_void_p),
                ("degree", C.c_int32), ("eig_ratio", C.c_double), ("eig_safety", C.c_double),
                ("inv_mu", C.c_double), ("mask0", C.c_void_p), ("nP1", C.c_int32),
                ("F", C.c_int32), ("rp11", C.c_void_p), ("ci11", C.c_void_p),
                ("nnz11", C.c_longlong), ("M1", C.c_void_p), ("St", C.c_void_p),
                ("dinvp", C.c_void_p), ("pmask", C.c_void_p),
                ("Pd", C.c_double * _lib.MAXF), ("Ps", C.c_double * _lib.MAXF),
                ("pdeg", C.c_int32), ("plmin", C.c_double * _lib.MAXF),
                ("plmax", C.c_double * _lib.MAXF)]


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
    L.mpet_solver_destroy.argtypes = [P]
    L.mpet_solver_destroy.restype = C.c_int
    L.mpet_solver_solve.argtypes = [P, P, P, P, P]
    L.mpet_solver_solve.restype = C.c_int
    L.mpet_solver_set_bnorm.argtypes = [P, C.c_double]
    L.mpet_solver_set_bnorm.restype = C.c_int
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
    outer_tol: float = 1e-8
    inner_rtol: float = 1e-4
    max_outer: int = 30
    check_every: int = 0              # 0: 1 for fp64, 10 for mixed
    use_graphs: bool = None           # default: mixed only
    use_csr: bool = None              # default: fp64 only
    amg: amg.AmgParams = field(default_factory=amg.AmgParams)
    p_reduction: float = 0.05         # Chebyshev error target of the pressure blocks
    max_pdeg: int = 12
    # weight of the total-pressure mass in the preconditioner.  The reference
    # form prec uses the bare mass p0 w0 dx (mpettotalpressuresolver.py:282),
    # which is off from the p0 Schur complement ~ (1/(2 mu) + 1/lambda) M by
    # a factor ~mu and costs ~4x MINRES iterations at E = 1500; "robust"
    # uses 1/(2 mu) + 1/lambda, "reference" the bare mass.
    p0_weight: str = "robust"


def _csr_i32(rp):
    if rp[-1] >= 2**31:
        raise ValueError("level too large for int32 row pointers")
    return rp.astype(np.int32)


def _pick_lpr(avg):
    for l in (4, 8, 16, 32):
        if avg <= 1.5 * l:
            return l
    return 64


class KrylovSolver:
    """Owns the device preconditioner data and the native solver handle."""

    def __init__(self, system, bc_dofs, config: SolverConfig = None):
        self.sys = system
        self.cfg = cfg = config or SolverConfig()
        dev = system.device
        self.keep = []                       # tensors the native handle borrows
        mixed = cfg.mode == "mixed"
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
        if cfg.p0_weight == "robust":
            self.Pd[0] = 0.5 / system.mu + 1.0 / system.lmbda
        elif cfg.p0_weight != "reference":
            raise ValueError("p0_weight must be 'robust' or 'reference'")
        self._build_hierarchy()
        self._build_pressure()
        L = _bind()
        self.prec64 = self._prec_desc(0) if not mixed else None
        self.prec32 = self._prec_desc(1) if mixed else None
        use_csr = cfg.use_csr if cfg.use_csr is not None else (not mixed)
        self.csr_desc = None
        if use_csr and not mixed:
            if system.csr is None:
                system.csr = system.expand_csr(0)
            rp, ci, v = system.csr
            avg = (system.nnz_csr if hasattr(system, "nnz_csr") else v.numel()) / system.n
            self.csr_desc = CsrDesc(system.n, ptr(rp), ptr(ci), ptr(v), _pick_lpr(avg))
        check_every = cfg.check_every or (1 if not mixed else 10)
        graphs = cfg.use_graphs if cfg.use_graphs is not None else mixed
        self.opt = SolverOptions(1 if mixed else 0, 1 if self.csr_desc is not None else 0,
                                 cfg.rtol, cfg.maxit, check_every, 1 if graphs else 0,
                                 cfg.outer_tol, cfg.inner_rtol, cfg.max_outer)
        self.op64 = system.op_desc(0)
        self.op32 = system.op_desc(1) if mixed else None
        h = C.c_void_p()
        check(L.mpet_solver_create(C.byref(self.op64),
                                   C.byref(self.op32) if self.op32 is not None else None,
                                   C.byref(self.csr_desc) if self.csr_desc is not None else None,
                                   C.byref(self.prec64) if self.prec64 is not None else None,
                                   C.byref(self.prec32) if self.prec32 is not None else None,
                                   C.byref(self.opt), C.byref(h)))
        self.handle = h
        self.last_stats = None

    # -- setup ----------------------------------------------------------------
    def _build_hierarchy(self):
        s = self.sys
        rp, ci = s.pat_host["22"]
        lap = s.prim_t["Lap"].cpu().numpy()
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

    def _level_arrays(self, dtype):
        if dtype in self.level_dev:
            return self.level_dev[dtype]
        s = self.sys
        dev = s.device
        tdt = torch.float64 if dtype == 0 else torch.float32
        out = []
        for l, lev in enumerate(self.levels):
            d = {}
            if l == 0:
                d["A_rp"], d["A_ci"] = s.pat_dev["22"]
                d["A_v"] = s.prim_t["Lap"] if dtype == 0 else s.prim_t["Lap"].to(tdt)
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
        cinv = torch.as_tensor(self.cinv).to(dev, tdt)
        self.level_dev[dtype] = (out, cinv)
        return out, cinv

    def _build_pressure(self):
        """Inverse diagonals and Chebyshev intervals of the pressure blocks
        (Pd_f M1 + Ps_f St_f, Dirichlet rows eliminated)."""
        s = self.sys
        rp, ci = s.pat_host["11"]
        M1 = s.prim_t["M1"].cpu().numpy()
        St = s.prim_t["St"].cpu().numpy().reshape(s.A, -1)
        n1 = s.nP1
        F = s.F
        Mm = sp.csr_matrix((M1, ci, rp), shape=(n1, n1))
        pmask = self.pmask.cpu().numpy().reshape(n1, F).astype(bool)
        dinv = np.ones((n1, F))
        lmin = np.zeros(F)
        lmax = np.zeros(F)
        dM = Mm.diagonal()
        Pdw = self.Pd
        for f in range(F):
            Pf = Pdw[f] * Mm
            dS = np.zeros(n1)
            if f >= 1:
                Sf = sp.csr_matrix((St[f - 1], ci, rp), shape=(n1, n1))
                Pf = Pf + s.Ps[f] * Sf
                dS = Sf.diagonal()
            free = ~pmask[:, f]
            keep = sp.diags(free.astype(np.float64))
            Pe = (keep @ Pf @ keep + sp.diags((~free).astype(np.float64))).tocsr()
            d = Pe.diagonal()
            dinv[:, f] = 1.0 / d
            # largest eigenvalue of D^-1 P: power iteration (D-inner product)
            x = np.random.default_rng(7).random(n1) * free
            lam = 1.0
            for _ in range(40):
                y = (Pe @ x) / d
                lam = float(x @ (d * y)) / float(x @ (d * x))
                x = y / np.linalg.norm(y)
            lmax[f] = 1.05 * lam
            # lower bound: P >= Pd M >= Pd/2 D_M  (P1 mass, Wathen) and
            # D = Pd D_M + Ps D_S  =>  lambda_min >= 1/2 min_i Pd D_M/(Pd D_M + Ps D_S)
            ratio = (Pdw[f] * dM) / (Pdw[f] * dM + (s.Ps[f] * dS if f >= 1 else 0.0))
            lmin[f] = 0.5 * float(ratio[free].min()) if free.any() else 1.0
        self.dinvp_host = dinv.reshape(-1)
        self.plmin, self.plmax = lmin, lmax
        # common degree reaching the error target in every field
        deg = 1
        for f in range(F):
            kappa = lmax[f] / lmin[f]
            rho = (np.sqrt(kappa) - 1) / (np.sqrt(kappa) + 1)
            k = 1
            while 2 * rho ** k / (1 + rho ** (2 * k)) > self.cfg.p_reduction and k < self.cfg.max_pdeg:
                k += 1
            deg = max(deg, k)
        self.pdeg = deg

    def _prec_desc(self, dtype):
        s = self.sys
        cfg = self.cfg
        levs, cinv = self._level_arrays(dtype)
        tdt = torch.float64 if dtype == 0 else torch.float32
        self.keep += [levs, cinv]
        d = PrecDesc()
        d.nlev = len(self.levels)
        if d.nlev > MAXLEV:
            raise ValueError("hierarchy too deep")
        for l, (lev, ld) in enumerate(zip(self.levels, levs)):
            L = d.lev[l]
            L.n = lev["n"]
            L.nc = lev["nc"]
            L.A_rp, L.A_ci, L.A_v = ptr(ld["A_rp"]), ptr(ld["A_ci"]), ptr(ld["A_v"])
            L.dinv = ptr(ld["dinv"])
            L.rho = lev["rho"]
            nnz = ld["A_ci"].numel()
            L.lpr = _pick_lpr(nnz / max(lev["n"], 1))
            if lev["nc"]:
                L.P_rp, L.P_ci, L.P_v = ptr(ld["P_rp"]), ptr(ld["P_ci"]), ptr(ld["P_v"])
                L.R_rp, L.R_ci, L.R_v = ptr(ld["R_rp"]), ptr(ld["R_ci"]), ptr(ld["R_v"])
                nnzp = ld["P_ci"].numel()
                L.lprP = _pick_lpr(nnzp / max(lev["n"], 1))
                L.lprR = _pick_lpr(nnzp / max(lev["nc"], 1))
        d.coarse_inv = ptr(cinv)
        d.degree = cfg.amg.smoother_degree
        d.eig_ratio = cfg.amg.eig_ratio
        d.eig_safety = cfg.amg.eig_safety
        d.inv_mu = 1.0 / s.mu
        d.mask0 = ptr(self.mask0)
        d.nP1 = s.nP1
        d.F = s.F
        rp11, ci11 = s.pat_dev["11"]
        d.rp11, d.ci11 = ptr(rp11), ptr(ci11)
        d.nnz11 = s.nnz["11"]
        vals = s.op64 if dtype == 0 else s.op32
        d.M1 = ptr(vals["M1"])
        d.St = ptr(vals["St"])
        dinvp = torch.as_tensor(self.dinvp_host).to(s.device, tdt)
        self.keep.append(dinvp)
        d.dinvp = ptr(dinvp)
        d.pmask = ptr(self.pmask)
        for f in range(s.F):
            d.Pd[f] = self.Pd[f]
            d.Ps[f] = s.Ps[f]
            d.plmin[f] = self.plmin[f]
            d.plmax[f] = self.plmax[f]
        d.pdeg = self.pdeg
        return d

    # -- use ------------------------------------------------------------------
    def solve(self, b, x, bnorm=0.0, raise_on_fail=True):
        """Solve one step: b full right-hand side (FP64 device), x initial
        guess with Dirichlet values in place (overwritten by the solution)."""
        L = _bind()
        st = SolveStats()
        if bnorm:
            L.mpet_solver_set_bnorm(self.handle, float(bnorm))
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
