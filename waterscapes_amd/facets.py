"""Exterior-facet terms of the total-pressure step on the GPU
(csrc/facets.hip): traction s (L0, :313), network fluxes I_i (L1_i, :322)
and Robin conditions beta_i, p_robin_i (L2_i, :324) of
mpet/mpettotalpressuresolver.py.

Expression coefficients are interpolated into P_k on each facet (their
declared degree, Expression semantics) and every integrand - a coefficient
times a test function, or a product of two interpolated coefficients times a
test function - is integrated exactly with reference-triangle tables (exact
rationals, refel.py).  The per-facet local vectors are computed on the
device and scattered node-major (one writer per entry, fixed order), so the
result is bitwise reproducible.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib, refel
from ._lib import check, ptr


def _bind():
    L = _lib.lib()
    P = C.c_void_p
    L.mpet_facet_local.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P, P, P,
                                   C.c_longlong, P, P, P, C.c_double, P, P]
    L.mpet_facet_scatter_vec.argtypes = [C.c_int, C.c_int, C.c_int, P, P, P, P, P, P]
    L.mpet_facet_scatter_mat.argtypes = [C.c_int, P, P, P, P, P, P, P, P, P]
    for f in (L.mpet_facet_local, L.mpet_facet_scatter_vec, L.mpet_facet_scatter_mat):
        f.restype = C.c_int
    return L


def facet_product_weights(k1, k2, test_degree=1):
    """T[a, m, n] = (1/|F|) int_F v_a phi^(k1)_m phi^(k2)_n on the reference
    triangle (v: P1 test functions, or P2 with test_degree 2)."""
    key = (k1, k2, test_degree)
    if key not in _TABLES:
        lam = refel.p1_basis(3)
        if test_degree == 2:
            test = [lam[i] * (2 * lam[i] - 1) for i in range(3)] + \
                   [4 * lam[i] * lam[j] for (i, j) in ((1, 2), (0, 2), (0, 1))]
        else:
            test = lam
        b1, b2 = refel.pk_basis(k1, 3), refel.pk_basis(k2, 3)
        _TABLES[key] = np.array([[[float((t * p * q).mean()) for q in b2] for p in b1]
                                 for t in test])
    return _TABLES[key]


_TABLES = {}


class FacetSet:
    """Exterior facets carrying one marker value (the measure ds(value))."""

    def __init__(self, space, markers, value, device):
        mesh = space.mesh
        dm = space.dofmap
        vals = np.asarray(markers.array() if hasattr(markers, "array") else markers)
        sel = np.nonzero((vals == value) & mesh.exterior)[0]
        self.n = int(sel.size)
        self.device = device
        cl = mesh.facet_cell[sel]
        lf = mesh.facet_local[sel]
        lv = np.array(refel.FACES)[lf]                                     # nF,3
        self.p1 = dm.cell_p1[cl[:, None], lv].astype(np.int64)             # P1 nodes
        l2 = np.array([refel.face_local_p2_nodes(f) for f in range(4)])[lf]
        self.p2 = dm.cell_p2[cl[:, None], l2].astype(np.int64)             # P2 nodes
        X = mesh.coordinates()
        cells = mesh.cells()
        self.xv = X[cells[cl[:, None], lv]]                                # nF,3,3
        cr = np.cross(self.xv[:, 1] - self.xv[:, 0], self.xv[:, 2] - self.xv[:, 0])
        a2 = np.linalg.norm(cr, axis=1)
        self.area = 0.5 * a2
        opp = X[cells[cl, lf]]                                            # opposite vertex
        nrm = cr / np.where(a2 > 0, a2, 1.0)[:, None]
        sgn = np.sign(np.einsum("fi,fi->f", nrm, self.xv[:, 0] - opp))
        self.normals = nrm * sgn[:, None]
        self.area_d = torch.as_tensor(self.area).to(device)
        self._inc = {}
        self._lat_dev = {}
        self._nrm_dev = {}

    def lattice(self, k):
        """(nF, nk, 3) coordinates of the P_k nodes of every facet."""
        bary = refel.pk_nodes_bary(k, 3)                                   # nk,3
        return np.einsum("mk,fkd->fmd", bary, self.xv)

    def incidence(self, which):
        """Node-major incidence of the facets' test nodes (which = 1: the 3
        P1 nodes, 2: the 6 P2 nodes): (rows, inc_rp, inc) with inc entries
        f * ntest + a sorted by (node, f, a)."""
        if which not in self._inc:
            nodes = self.p1 if which == 1 else self.p2
            ntest = nodes.shape[1]
            flat = nodes.reshape(-1)
            order = np.lexsort((np.arange(flat.size), flat))              # by node, then (f, a)
            rows, counts = np.unique(flat[order], return_counts=True)
            rp = np.zeros(rows.size + 1, dtype=np.int64)
            np.cumsum(counts, out=rp[1:])
            dev = self.device
            self._inc[which] = (rows, torch.as_tensor(rp.astype(np.int32)).to(dev),
                                torch.as_tensor(order.astype(np.int32)).to(dev), ntest)
        return self._inc[which]

    def eval(self, expr, t=None):
        """Nodal values (nF, nk, ncomp), a device tensor, of an Expression
        interpolated into P_degree on every facet (normals passed to
        traction expressions).  Evaluated where the facet lattices live
        (Expression.eval_device); host-only coefficients are copied."""
        k = int(getattr(expr, "degree", 0))
        if k not in self._lat_dev:
            Xf = self.lattice(k)
            self._lat_dev[k] = (torch.as_tensor(Xf.reshape(-1, 3)).to(self.device), Xf.shape[1])
        Xd, nk = self._lat_dev[k]
        nF = self.n
        if hasattr(expr, "eval_normal"):
            if self._nrm_dev.get(k) is None:
                self._nrm_dev[k] = torch.as_tensor(np.repeat(self.normals, nk, axis=0)).to(
                    self.device)
            v = expr.eval_normal_device(Xd, self._nrm_dev[k], t)
        elif hasattr(expr, "eval_device"):
            v = expr.eval_device(Xd, t)
        else:
            v = torch.as_tensor(np.asarray(expr.eval(Xd.cpu().numpy(), t),
                                           dtype=np.float64)).to(self.device)
        v = v.to(torch.float64)
        if v.dim() == 0:
            v = v.expand(nF * nk)
        v = v.reshape(nF * nk, -1)
        return v.contiguous().reshape(nF, nk, v.shape[1]), k


class FacetTerm:
    """scale * int_{facets} (coefficient product) . test, added into b.

    kind "source": one coefficient u (ncomp components) of degree k:
        loc[f][a][c] = scale |F| sum_m Q[a][m] u[f][m][c]
    kind "product": int beta_h q_h w with two scalar coefficients;
    kind "robin_prev": int beta_h x_h w with x the P1 field of a vector
        (gathered at the facet's DOFs each call)."""

    def __init__(self, fs: FacetSet, test_degree, dof_base, dof_stride, dof_off, ncomp):
        self.fs = fs
        self.test_degree = test_degree
        rows, rp, inc, ntest = fs.incidence(test_degree)
        self.rows = rows
        self.inc_rp, self.inc, self.ntest = rp, inc, ntest
        self.ncomp = ncomp
        row_dof = dof_base + dof_stride * rows + dof_off
        self.row_dof = torch.as_tensor(row_dof.astype(np.int64)).to(fs.device)
        self.loc = torch.empty(max(fs.n, 1) * ntest * ncomp, dtype=torch.float64,
                               device=fs.device)
        self._tables = {}

    def _scatter(self, b):
        check(_bind().mpet_facet_scatter_vec(int(self.rows.size), self.ntest, self.ncomp,
                                             ptr(self.row_dof), ptr(self.inc_rp),
                                             ptr(self.inc), ptr(self.loc), ptr(b),
                                             _lib.stream_ptr()))

    def add_source(self, b, expr, scale, t=None):
        fs = self.fs
        if fs.n == 0:
            return
        u, k = fs.eval(expr, t)
        if u.shape[2] != self.ncomp:
            raise ValueError(f"facet coefficient has {u.shape[2]} components, "
                             f"expected {self.ncomp}")
        T = np.asarray(refel.facet_source_weights(self.test_degree, k))[:, :, None]
        self._local(T, u, None, None, scale)
        self._scatter(b)

    def add_product(self, b, expr_u, expr_w, scale, t=None):
        fs = self.fs
        if fs.n == 0:
            return
        u, ku = fs.eval(expr_u, t)
        w, kw = fs.eval(expr_w, t)
        T = facet_product_weights(ku, kw, self.test_degree)
        self._local(T, u, w[:, :, 0].contiguous(), None, scale)
        self._scatter(b)

    def add_robin_prev(self, b, expr_beta, x, gdof, scale, t=None):
        """b += scale int beta_h x_h w, x_h the P1 interpolant of the vector
        x at the facet DOFs gdof (nF, 3)."""
        fs = self.fs
        if fs.n == 0:
            return
        w, kw = fs.eval(expr_beta, t)
        T = facet_product_weights(1, kw, self.test_degree)
        self._local(T, None, w[:, :, 0].contiguous(), (x, gdof), scale)
        self._scatter(b)

    def _local(self, T, u, w, gather, scale):
        fs = self.fs
        nm, nn = T.shape[1], T.shape[2]
        Th = np.ascontiguousarray(T, dtype=np.float64)
        key = (Th.shape, Th.tobytes())           # reference tables: uploaded once
        Td = self._tables.get(key)
        if Td is None:
            Td = self._tables[key] = torch.as_tensor(Th).to(fs.device)
        self._keep = (Td, u, w)
        if gather is not None:
            x, gdof = gather
            rc = _bind().mpet_facet_local(fs.n, self.ntest, nm, nn, 1, ptr(Td), ptr(fs.area_d),
                                          None, 0, ptr(x), ptr(gdof), ptr(w), float(scale),
                                          ptr(self.loc), _lib.stream_ptr())
        else:
            rc = _bind().mpet_facet_local(fs.n, self.ntest, nm, nn, self.ncomp, ptr(Td),
                                          ptr(fs.area_d), ptr(u), nm * self.ncomp, None, None,
                                          ptr(w), float(scale), ptr(self.loc),
                                          _lib.stream_ptr())
        check(rc)


def robin_matrix(fs: FacetSet, expr_beta, t, pat_dev, vals, scale=1.0):
    """vals (a P1-pattern array on the device) += scale int_F beta_h psi_a psi_b
    over the facets of fs (beta at time t)."""
    if fs.n == 0:
        return
    dev = fs.device
    w, kw = fs.eval(expr_beta, t)
    T = facet_product_weights(1, kw, 1)                     # a, b, n
    Td = torch.as_tensor(np.ascontiguousarray(T)).to(dev)
    eye = torch.eye(3, dtype=torch.float64, device=dev).reshape(-1)   # u[m][c] = delta
    loc = torch.empty(fs.n * 9, dtype=torch.float64, device=dev)
    wd = w[:, :, 0].contiguous()
    L = _bind()
    check(L.mpet_facet_local(fs.n, 3, 3, T.shape[2], 3, ptr(Td), ptr(fs.area_d), ptr(eye), 0,
                             None, None, ptr(wd), float(scale), ptr(loc), _lib.stream_ptr()))
    rows, inc_rp, inc, _ = fs.incidence(1)
    rows_d = torch.as_tensor(rows.astype(np.int32)).to(dev)
    fnode = torch.as_tensor(fs.p1.astype(np.int32).reshape(-1)).to(dev)
    rp11, ci11 = pat_dev
    check(L.mpet_facet_scatter_mat(int(rows.size), ptr(rows_d), ptr(inc_rp), ptr(inc),
                                   ptr(fnode), ptr(rp11), ptr(ci11), ptr(loc), ptr(vals),
                                   _lib.stream_ptr()))
    torch.cuda.current_stream().synchronize()
