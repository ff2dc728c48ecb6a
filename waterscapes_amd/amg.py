"""Smoothed-aggregation AMG hierarchy for the displacement block.

Setup runs natively (csrc/amg_setup.cpp) on the host; this module moves the
hierarchy into numpy arrays (host) and then into HBM (device.AmgDevice).
The V-cycle itself runs on the GPU (csrc/krylov.hip).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib


class AmgParamsC(C.Structure):
    _fields_ = [("strength", C.c_double), ("max_coarse", C.c_int), ("max_levels", C.c_int),
                ("power_iters", C.c_int), ("omega_scale", C.c_double)]


@dataclass
class AmgParams:
    strength: float = 0.08
    max_coarse: int = 400
    max_levels: int = 12
    power_iters: int = 30
    omega_scale: float = 4.0 / 3.0
    smoother_degree: int = 2          # Chebyshev degree (pre and post)
    level_degrees: tuple = ()         # degrees of levels 0, 1, .. (0 / absent: smoother_degree)

    def degree(self, level):
        d = self.level_degrees[level] if level < len(self.level_degrees) else 0
        return int(d) if d else int(self.smoother_degree)
    eig_ratio: float = 8.0            # Chebyshev interval [rho/ratio, 1.1 rho]
    eig_safety: float = 1.1


def _bind():
    L = _lib.lib()
    P = C.c_void_p
    L.mpet_amg_setup.argtypes = [C.c_int, P, P, P, P, C.c_int, P, P, P, C.POINTER(C.c_void_p)]
    L.mpet_amg_setup.restype = C.c_int
    L.mpet_amg_num_levels.argtypes = [P]
    L.mpet_amg_num_levels.restype = C.c_int
    L.mpet_amg_level_sizes.argtypes = [P, C.c_int, P]
    L.mpet_amg_level_sizes.restype = C.c_int
    L.mpet_amg_level_copy.argtypes = [P, C.c_int] + [P] * 11
    L.mpet_amg_level_copy.restype = C.c_int
    L.mpet_amg_coarse_inverse.argtypes = [P, P]
    L.mpet_amg_coarse_inverse.restype = C.c_int
    L.mpet_amg_destroy.argtypes = [P]
    L.mpet_amg_destroy.restype = None
    L.mpet_amg_set_threads.argtypes = [C.c_int]
    L.mpet_amg_set_threads.restype = C.c_int
    return L


def set_host_threads(n):
    """OpenMP threads of the hierarchy setups run next by this thread."""
    _lib.check(_bind().mpet_amg_set_threads(int(n)))


def setup_host(rowptr, cols, vals, params: AmgParams = None, P0=None):
    """Build the hierarchy for the SPD scalar CSR matrix (host arrays); P0 an
    optional prescribed first prolongator (scipy CSR, n x n1).
    Returns a list of level dicts (A_rp, A_ci, A_v, dinv, rho, P_*, R_*) and
    the dense inverse of the coarsest operator."""
    params = params or AmgParams()
    L = _bind()
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    n = rowptr.size - 1
    prm = AmgParamsC(params.strength, params.max_coarse, params.max_levels,
                     params.power_iters, params.omega_scale)
    h = C.c_void_p()
    if P0 is not None:
        p_rp = np.ascontiguousarray(P0.indptr, dtype=np.int64)
        p_ci = np.ascontiguousarray(P0.indices, dtype=np.int32)
        p_v = np.ascontiguousarray(P0.data, dtype=np.float64)
        p0args = (P0.shape[1], p_rp.ctypes.data, p_ci.ctypes.data, p_v.ctypes.data)
    else:
        p0args = (0, None, None, None)
    _lib.check(L.mpet_amg_setup(n, rowptr.ctypes.data, cols.ctypes.data, vals.ctypes.data,
                                C.byref(prm), *p0args, C.byref(h)))
    try:
        nl = L.mpet_amg_num_levels(h)
        levels = []
        for l in range(nl):
            sz = np.zeros(4, dtype=np.int64)
            L.mpet_amg_level_sizes(h, l, sz.ctypes.data)
            nr, nnz, nc, nnzp = (int(s) for s in sz)
            lev = dict(n=nr, A_rp=np.empty(nr + 1, np.int64), A_ci=np.empty(nnz, np.int32),
                       A_v=np.empty(nnz, np.float64), dinv=np.empty(nr, np.float64),
                       rho=np.zeros(1), nc=nc)
            ptrs = [lev["A_rp"], lev["A_ci"], lev["A_v"], lev["dinv"], lev["rho"]]
            if nc:
                lev.update(P_rp=np.empty(nr + 1, np.int64), P_ci=np.empty(nnzp, np.int32),
                           P_v=np.empty(nnzp, np.float64), R_rp=np.empty(nc + 1, np.int64),
                           R_ci=np.empty(nnzp, np.int32), R_v=np.empty(nnzp, np.float64))
                ptrs += [lev[k] for k in ("P_rp", "P_ci", "P_v", "R_rp", "R_ci", "R_v")]
            else:
                ptrs += [None] * 6
            L.mpet_amg_level_copy(h, l, *[p.ctypes.data if p is not None else None for p in ptrs])
            lev["rho"] = float(lev["rho"][0])
            levels.append(lev)
        ncoarse = levels[-1]["n"]
        cinv = np.empty(ncoarse * ncoarse, dtype=np.float64)
        L.mpet_amg_coarse_inverse(h, cinv.ctypes.data)
    finally:
        L.mpet_amg_destroy(h)
    return levels, cinv.reshape(ncoarse, ncoarse)


def chebyshev_coefficients(lmin, lmax, degree):
    """Coefficients of the Chebyshev iteration on D^-1 A over [lmin, lmax]
    (Saad, Iterative Methods, Alg. 12.1): step 0 is d = r / theta; step i
    (i >= 1) is d = cd[i] d + cr[i] D^-1 (b - A x).  Returns (cd, cr)."""
    theta = 0.5 * (lmax + lmin)
    delta = 0.5 * (lmax - lmin)
    sigma = theta / delta
    rho = 1.0 / sigma
    cd = [0.0]
    cr = [1.0 / theta]
    for _ in range(1, degree):
        rho_new = 1.0 / (2.0 * sigma - rho)
        cd.append(rho_new * rho)
        cr.append(2.0 * rho_new / delta)
        rho = rho_new
    return np.array(cd), np.array(cr)


def p1_embedding(space, dirichlet_p2_nodes):
    """Prolongator from P1 (free vertices) into P2: vertex nodes copy the
    vertex value, edge nodes average their two endpoints (exact: P1 is a
    subspace of P2).  Dirichlet vertices carry no coarse unknown and
    Dirichlet P2 nodes get empty rows, so the coarse space is P1 with the
    same homogeneous Dirichlet condition."""
    import scipy.sparse as sp
    from . import refel
    dm = space.dofmap
    c1 = dm.cell_p1.astype(np.int64)
    c2 = dm.cell_p2.astype(np.int64)
    nV = dm.nP1
    isD = np.zeros(dm.nP2, dtype=bool)
    isD[np.asarray(dirichlet_p2_nodes, dtype=np.int64)] = True
    vert_node = np.empty(nV, dtype=np.int64)
    for k in range(4):
        vert_node[c1[:, k]] = c2[:, k]
    vD = isD[vert_node]
    cmap = -np.ones(nV, dtype=np.int64)
    cmap[~vD] = np.arange(int((~vD).sum()))
    rows = [vert_node[~vD]]
    cols = [cmap[~vD]]
    vals = [np.ones(int((~vD).sum()))]
    en = np.concatenate([c2[:, 4 + m] for m in range(6)])
    ea = np.concatenate([c1[:, i] for (i, j) in refel.EDGES])
    eb = np.concatenate([c1[:, j] for (i, j) in refel.EDGES])
    _, idx = np.unique(en, return_index=True)
    en, ea, eb = en[idx], ea[idx], eb[idx]
    for a in (ea, eb):
        ok = (~isD[en]) & (~vD[a])
        rows.append(en[ok])
        cols.append(cmap[a[ok]])
        vals.append(np.full(int(ok.sum()), 0.5))
    P0 = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                       shape=(dm.nP2, int((~vD).sum()))).tocsr()
    P0.sort_indices()
    return P0
