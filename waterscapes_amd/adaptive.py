"""Error indicators and estimators of AdaptiveMPETSolver
(mpet/adaptivempetsolver.py:11-225) for the two-field MPETSolver, computed on
the GPU (csrc/indicators.hip, `mpet_error_indicators`).

`ErrorEstimator(problem, solver, f_)` holds what
AdaptiveMPETSolver.__init_error_estimators (:82-163) builds: it evaluates the
DG0 cell indicators zeta_p, zeta_u, zeta_u_dt (cell residual times h^2 plus
half of the interior-facet jump terms of the cell's facets) and the global
R4p for the solver's current (up) and previous (up_) states.  The
bookkeeping methods estimate_error_at_t, add_to_indicators,
pop_error_estimators, compute_error_indicators and compute_error_estimate
(:165-222) and the markers maximal_mark / dorfler_mark (:11-64) follow the
reference; indicator arrays are numpy arrays of one value per cell (the
reference's DG0 vectors).  The time-adaptive loop and mesh refinement of
AdaptiveMPETSolver are not part of this module.

Material parameters must be constants (the reference wraps them in
Constant, :106-111); f, f_ and g are Expressions (P_k interpolation per
cell, as in the forms).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib, forms, refel
from ._lib import check, ptr

MAXF = _lib.MAXF


class IndDesc(C.Structure):
    _fields_ = [("nC", C.c_int32), ("nP2", C.c_int32), ("nP1", C.c_int32), ("A", C.c_int32),
                ("coords", C.c_void_p), ("cells", C.c_void_p), ("cell_p2", C.c_void_p),
                ("mu", C.c_double), ("lmbda", C.c_double), ("dt", C.c_double),
                ("alpha", C.c_double * MAXF), ("c", C.c_double * MAXF),
                ("K", C.c_double * MAXF), ("S", C.c_double * (MAXF * MAXF)),
                ("nq", C.c_int32), ("qbary", C.c_void_p), ("qw", C.c_void_p),
                ("nkf", C.c_int32), ("nkg", C.c_int32), ("Lf", C.c_void_p), ("Lg", C.c_void_p),
                ("f", C.c_void_p), ("fp", C.c_void_p), ("g", C.c_void_p)]


def maximal_mark(indicators, fraction):
    """Mark the ceil(fraction * nC) cells with the largest indicators
    (:11-34).  Returns a boolean array over cells."""
    etas = np.asarray(indicators, dtype=np.float64)
    idx = np.argsort(-etas)
    n = math.ceil(fraction * etas.size)
    out = np.zeros(etas.size, dtype=bool)
    out[idx[:n]] = True
    return out


def dorfler_mark(indicators, fraction):
    """Doerfler marking (:36-64): mark the largest indicators until their sum
    reaches fraction * sum."""
    etas = np.asarray(indicators, dtype=np.float64)
    idx = np.argsort(-etas)
    stop = fraction * etas.sum()
    out = np.zeros(etas.size, dtype=bool)
    eta = 0.0
    for i in idx:
        if eta >= stop:
            break
        out[i] = True
        eta += etas[i]
    return out


def interior_facets(mesh):
    """(nF x 4 int32: cell a, local face a, cell b, local face b) for the
    interior facets, and (nC x 4) the interior facet index of each local face
    (-1 on the boundary)."""
    cf = mesh.cell_facets.reshape(-1)
    interior = ~mesh.exterior
    fid = np.full(mesh.num_facets(), -1, dtype=np.int64)
    fid[interior] = np.arange(int(interior.sum()))
    slots = np.nonzero(interior[cf])[0]
    order = np.argsort(cf[slots], kind="stable")
    s = slots[order].reshape(-1, 2)
    fac = np.stack([s[:, 0] // 4, s[:, 0] % 4, s[:, 1] // 4, s[:, 1] % 4], axis=1)
    cell_f = fid[cf].reshape(-1, 4)
    return np.ascontiguousarray(fac, dtype=np.int32), np.ascontiguousarray(cell_f, dtype=np.int32)


def _const(v, name):
    a = np.asarray(v, dtype=np.float64)
    if a.ndim > 0 and a.size != 1:
        raise NotImplementedError(f"{name}: the error indicators take constant parameters")
    return float(a)


class ErrorEstimator:
    """Indicators of mpet/adaptivempetsolver.py:82-222 for a two-field
    MPETSolver (see the module docstring)."""

    def __init__(self, problem, solver, f_):
        if not torch.cuda.is_available():
            raise RuntimeError("ErrorEstimator needs a GPU (no CPU fallback)")
        space = solver.up.function_space()
        if space.total_pressure or space.nmult:
            raise NotImplementedError("error indicators of the two-field formulation without "
                                      "nullspace multipliers")
        self.problem = problem
        self.solver = solver
        self.f_ = f_
        dev = torch.device(solver.params["device"])
        self.device = dev
        mesh = space.mesh
        dm = space.dofmap
        self.space = space
        P = problem.params
        A = int(P["J"])
        self.A = A
        mu, lmbda = forms.lame(float(P["E"]), float(P["nu"]))
        d = IndDesc()
        d.nC, d.nP2, d.nP1, d.A = mesh.num_cells(), dm.nP2, dm.nP1, A
        self.coords = torch.as_tensor(mesh.coordinates()).to(dev)
        self.cells = torch.as_tensor(dm.cell_p1.astype(np.int32)).to(dev)
        self.cell_p2 = torch.as_tensor(dm.cell_p2.astype(np.int32)).to(dev)
        d.coords, d.cells, d.cell_p2 = ptr(self.coords), ptr(self.cells), ptr(self.cell_p2)
        d.mu, d.lmbda = mu, lmbda
        for j in range(A):
            d.alpha[j] = _const(P["alpha"][j], "alpha")
            d.c[j] = _const(P["c"][j], "c")
            d.K[j] = _const(P["K"][j], "K")
            for i in range(A):
                d.S[i * MAXF + j] = _const(P["S"][i][j], "S")
        # sources: P_k interpolation per cell
        self.kf = max(int(getattr(problem.f, "degree", 0)), int(getattr(f_, "degree", 0)))
        self.kg = max([int(getattr(g, "degree", 0)) for g in problem.g] + [0])
        qdeg = 2 * max(self.kf, self.kg, 1)
        qb, qw = refel.gm_rule(qdeg)
        self.qbary = torch.as_tensor(qb).to(dev)
        self.qw = torch.as_tensor(qw).to(dev)
        d.nq, d.qbary, d.qw = len(qw), ptr(self.qbary), ptr(self.qw)
        x = mesh.coordinates()
        c = dm.cell_p1
        self._X = {}
        self._L = {}
        degs = {self.kf, self.kg, int(getattr(problem.f, "degree", 0)),
                int(getattr(f_, "degree", 0))} | {int(getattr(g, "degree", 0)) for g in problem.g}
        for k in degs:
            bary = refel.pk_nodes_bary(k)
            self._X[k] = np.einsum("nk,ckd->cnd", bary, x[c]).reshape(-1, 3)
        for k in {self.kf, self.kg}:
            L = np.stack([refel.bpoly_eval(p, qb) for p in refel.pk_basis(k)], axis=1)
            self._L[k] = torch.as_tensor(np.ascontiguousarray(L)).to(dev)
        d.nkf, d.nkg = self._L[self.kf].shape[1], self._L[self.kg].shape[1]
        d.Lf, d.Lg = ptr(self._L[self.kf]), ptr(self._L[self.kg])
        self.desc = d
        fac, cell_f = interior_facets(mesh)
        self.nF = fac.shape[0]
        self.fac = torch.as_tensor(fac).to(dev)
        self.cell_f = torch.as_tensor(cell_f).to(dev)
        self.tmp = torch.empty(max(3 * self.nF, 1), dtype=torch.float64, device=dev)
        self.out = torch.empty(4 * d.nC, dtype=torch.float64, device=dev)
        self._cache = {}               # device node values of time-independent sources
        # lists of error estimate terms, one for each time (:155-159)
        self.eta_1s, self.eta_2s, self.eta_3s, self.eta_4s = [], [], [], []
        nC = d.nC
        self.eta_1Ks = np.zeros(nC)
        self.eta_2Ks = np.zeros(nC)
        self.eta_3Ks = np.zeros(nC)

    def _nodal(self, expr, k, ncomp):
        """(nC, nk, ncomp) values of expr's P_k' interpolant (k' its own
        degree) at the P_k nodes of every cell (k >= k': exact)."""
        ko = int(getattr(expr, "degree", 0))
        Xo = self._X[ko]
        nC = self.desc.nC
        v = np.asarray(expr.eval(Xo, None), dtype=np.float64)
        v = np.broadcast_to(v, (Xo.shape[0], ncomp) if ncomp > 1 else (Xo.shape[0],))
        v = v.reshape(nC, -1, ncomp)
        if ko != k:
            T = np.stack([refel.bpoly_eval(p, refel.pk_nodes_bary(k))
                          for p in refel.pk_basis(ko)], axis=1)          # nk, nko
            v = np.einsum("nm,cmj->cnj", T, v)
        return v

    def _values(self, slot, expr, k, ncomp):
        """Device P_k nodal values of expr on every cell; kept between calls
        for an expression that does not depend on time."""
        if expr is None or expr.is_zero():
            return None
        key = (slot, k, ncomp)
        dep = getattr(expr, "depends_on_time", lambda: True)()
        if not dep:
            hit = self._cache.get(key)
            # the entry keeps its expression: an identity check, not an id()
            # that a later object may reuse
            if hit is not None and hit[0] is expr:
                return hit[1]
        v = self._nodal(expr, k, ncomp)
        t = torch.as_tensor(np.array(v.reshape(-1), dtype=np.float64)).to(self.device)
        if not dep:
            # one entry per slot (f or f_): replacing problem.f or f_ evicts
            # the old expression's values
            self._cache[key] = (expr, t)
        return t

    def descriptor(self):
        """The kernel descriptor with the current dt and source values set,
        and the device tensors it points to (keep them alive while the
        descriptor is used)."""
        d = self.desc
        d.dt = float(self.solver.params["dt"])
        f = self._values("f", self.problem.f, self.kf, 3)
        fp = self._values("f_", self.f_, self.kf, 3)
        g = None
        if any(not gi.is_zero() for gi in self.problem.g):
            cols = [self._nodal(gi, self.kg, 1)[:, :, 0] for gi in self.problem.g]
            g = torch.as_tensor(np.ascontiguousarray(np.stack(cols, axis=2))).to(self.device)
        d.f = ptr(f) if f is not None else None
        d.fp = ptr(fp) if fp is not None else None
        d.g = ptr(g) if g is not None else None
        return d, (f, fp, g)

    def indicators(self):
        """(zeta_p, zeta_u, zeta_u_dt, r4) per cell (numpy) for the solver's
        current up and previous up_."""
        d, keep = self.descriptor()
        x = self.solver.up.data
        xp = self.solver.up_.data
        check(_lib.lib().mpet_error_indicators(C.byref(d), ptr(x), ptr(xp), ptr(self.fac),
                                               self.nF, ptr(self.cell_f), ptr(self.tmp),
                                               ptr(self.out), _lib.stream_ptr()))
        out = self.out.view(4, d.nC).cpu().numpy().copy()     # synchronises
        return out[0], out[1], out[2], out[3]

    # -- reference bookkeeping (:165-222) --------------------------------------
    def estimate_error_at_t(self):
        eta_p_K_m, eta_u_K_m, eta_u_dt_K_m, r4 = self.indicators()
        eta_p_m = eta_p_K_m.sum()
        eta_u_m = eta_u_K_m.sum()
        eta_u_dt_m = eta_u_dt_K_m.sum()
        eta_dt_p_m = r4.sum()
        tau_m = float(self.solver.params["dt"])
        self.eta_1s += [tau_m * eta_p_m]
        self.eta_2s += [eta_u_m]
        self.eta_3s += [tau_m * np.sqrt(eta_u_dt_m)]
        self.eta_4s += [tau_m * eta_dt_p_m]
        return (eta_p_K_m, eta_u_K_m, eta_u_dt_K_m, tau_m)

    def add_to_indicators(self, eta_Ks):
        (eta_p_K_m, eta_u_K_m, eta_u_dt_K_m, tau_m) = eta_Ks
        self.eta_1Ks += tau_m * eta_p_K_m
        self.eta_3Ks = tau_m * np.sqrt(eta_u_dt_K_m)
        self.eta_2Ks = np.maximum(self.eta_2Ks, eta_u_K_m)

    def pop_error_estimators(self):
        self.eta_1s.pop()
        self.eta_2s.pop()
        self.eta_3s.pop()
        self.eta_4s.pop()

    def compute_error_indicators(self):
        return self.eta_2Ks + self.eta_3Ks + np.sqrt(self.eta_1Ks)

    def compute_error_estimate(self):
        eta_1 = np.sqrt(np.sum(self.eta_1s))
        eta_2 = np.sqrt(np.max(self.eta_2s))
        eta_3 = np.sum(self.eta_3s)
        eta_4 = np.sqrt(np.sum(self.eta_4s))
        etas = (eta_1, eta_2, eta_3, eta_4)
        return (sum(etas), etas)
