"""Device-resident MPET system of one instance: mesh tables, assembled
operators (FP64 primitives, structured FP64/FP32 copies, optional scalar
CSR), Dirichlet masks and the per-step right-hand side.

Everything lives in HBM as torch tensors; kernels are launched through the
C ABI (include/mpet_hip.h) on torch's current stream.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr
from .fem import MixedSpace, cell_colouring
from . import forms


def _dev(a, dtype, device):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(device)


def block_partial_tables(cell_nodes, nnode, B):
    """Block partials of the cell-wise operator for a cell -> node table
    (nC x k tensor, any device): cells are taken in blocks of B; every
    (node, block) pair is one partial.  Partials are stored block-major:
    block b's are q in [bq[b], bq[b+1]) (ordered by node), partial q sums
    the local slots lsl[lrp[q] .. lrp[q+1]) (c * k + j, ascending; c the
    cell in the block).  Entries list the partials in the order (node,
    block), a node's contiguous from rp[node]: entry e holds partial pe[e],
    and partial q sits at entry gp[q].  Also returns the node table in
    structure-of-arrays layout and the number of partials."""
    i32, i64 = torch.int32, torch.int64
    dev = cell_nodes.device
    nC, k = cell_nodes.shape
    nodes = cell_nodes.reshape(-1).to(i64)              # slot e*k + j
    ns = nodes.numel()
    slot = torch.arange(ns, device=dev, dtype=i64)
    blk = slot // (B * k)
    nb = (nC + B - 1) // B
    key = nodes * nb + blk                              # partial (node, block)
    order = torch.argsort(key, stable=True)
    ks = key[order]
    first = torch.ones(ns, dtype=torch.bool, device=dev)
    first[1:] = ks[1:] != ks[:-1]
    pid_sorted = torch.cumsum(first.to(i64), 0) - 1
    npart = int(pid_sorted[-1].item()) + 1
    pid = torch.empty_like(pid_sorted)
    pid[order] = pid_sorted                             # entry of each slot
    pnode = nodes[order][first]
    rp = torch.zeros(nnode + 1, dtype=i64, device=dev)
    rp[1:] = torch.cumsum(torch.bincount(pnode, minlength=nnode), 0)
    pblk = blk[order][first]                            # block of each partial
    gp = torch.argsort(pblk, stable=True)               # block-major -> entry
    bq = torch.zeros(nb + 1, dtype=i64, device=dev)
    bq[1:] = torch.cumsum(torch.bincount(pblk, minlength=nb), 0)
    order2 = torch.argsort(blk * npart + pid, stable=True)   # by (block, entry, slot)
    lsl = (order2 - blk[order2] * (B * k)).to(torch.int16)
    cnt = torch.bincount(pid, minlength=npart)[gp]
    lrp = torch.zeros(npart + 1, dtype=i64, device=dev)
    lrp[1:] = torch.cumsum(cnt, 0)
    soa = nodes.reshape(nC, k).t().contiguous().to(i32)
    pe = torch.empty_like(gp)
    pe[gp] = torch.arange(npart, device=dev, dtype=i64)
    return dict(node=soa, rp=rp.to(i32), npart=npart, bq=bq.to(i32), lrp=lrp.to(i32),
                gp=gp.to(i32), lsl=lsl.contiguous(), pe=pe.to(i32))


def assemble_batch(systems):
    """Assemble the primitives of many systems created with assemble=False
    in one launch per colour index (mpet_assemble_primitives_batched), then
    finish each system's set-up."""
    if not systems:
        return
    n = len(systems)
    L = _lib.lib()
    meshes = (_lib.MeshDesc * n)(*[s.mesh_desc for s in systems])
    pats = (_lib.Patterns * n)(*[s.pat for s in systems])
    kc = (C.c_void_p * n)(*[ptr(s.Kcell) for s in systems])
    A = (C.c_int32 * n)(*[s.A for s in systems])
    mu = (C.c_double * n)(*[s.mu for s in systems])
    outs = (_lib.Primitives * n)(*[s.prim for s in systems])
    check(L.mpet_assemble_primitives_batched(n, meshes, pats, kc, A, mu, outs,
                                             _lib.stream_ptr()))
    for s in systems:
        s.finish_assembly()


class DeviceSystem:
    """Assembled total-pressure system of one MPET instance on one GPU."""

    def __init__(self, space: MixedSpace, params, dt, theta, K_cells=None, device="cuda",
                 build_csr=False, need_fp32=True, structured=True, robin=None, assemble=True):
        if not torch.cuda.is_available():
            raise RuntimeError("DeviceSystem needs a GPU (no CPU fallback)")
        self.device = torch.device(device)
        _lib.ensure_tables(self.device.index or 0)
        self.space = space
        self.A = A = space.A
        self.F = A + 1
        self.nP2, self.nP1 = space.nP2, space.nP1
        self.n = space.ndofs
        self.nu = space.nu
        mesh = space.mesh
        dm = space.dofmap
        self.mu, self.lmbda = forms.lame(params["E"], params["nu"])
        alpha, c, S = params["alpha"], params["c"], params["S"]
        self.dt, self.theta = float(dt), float(theta)
        self.Cm, self.Cs = forms.matrix_coefficients(alpha, c, S, self.lmbda, dt, theta)
        self.Bm, self.Bs = forms.rhs_coefficients(alpha, c, S, self.lmbda, dt, theta)
        self.Pd, self.Ps = forms.preconditioner_coefficients(alpha, c, S, self.lmbda, dt, theta)
        dev = self.device
        i32, f64 = torch.int32, torch.float64
        # mesh tables
        self.coords = _dev(mesh.coordinates(), f64, dev)
        self.cells = _dev(dm.cell_p1, i32, dev)
        self.cell_p2 = _dev(dm.cell_p2, i32, dev)
        perm, offs = cell_colouring(mesh)
        self.colour_cells = _dev(perm, i32, dev)
        self.colour_offsets = np.ascontiguousarray(offs, dtype=np.int64)
        if K_cells is None:
            K_cells = np.tile(np.asarray([float(k) for k in params["K"]]), (mesh.num_cells(), 1))
        self.Kcell = _dev(np.asarray(K_cells, dtype=np.float64).reshape(mesh.num_cells(), A),
                          f64, dev)
        self.mesh_desc = _lib.MeshDesc(
            mesh.num_vertices(), mesh.num_cells(), self.nP2, self.nP1, ptr(self.coords),
            ptr(self.cells), ptr(self.cell_p2), ptr(self.colour_cells),
            self.colour_offsets.ctypes.data, len(offs) - 1)
        # node patterns
        self.pat_host = {}
        self.pat_dev = {}
        for kind in ("22", "21", "12", "11"):
            rp, ci = dm.pattern(kind)
            if rp[-1] >= 2**31:
                raise ValueError("pattern too large for int32 indexing")
            self.pat_host[kind] = (rp, ci)
            self.pat_dev[kind] = (_dev(rp.astype(np.int32), i32, dev), _dev(ci, i32, dev))
        p = self.pat_dev
        self.pat = _lib.Patterns(ptr(p["22"][0]), ptr(p["22"][1]), ptr(p["21"][0]),
                                 ptr(p["21"][1]), ptr(p["12"][0]), ptr(p["12"][1]),
                                 ptr(p["11"][0]), ptr(p["11"][1]),
                                 int(self.pat_host["22"][1].size), int(self.pat_host["21"][1].size),
                                 int(self.pat_host["12"][1].size), int(self.pat_host["11"][1].size))
        self.nnz = {k: int(v[1].size) for k, v in self.pat_host.items()}
        self.dmask = torch.zeros(self.n, dtype=torch.uint8, device=dev)
        self._alloc_primitives()
        self._rest = (robin or [], structured, need_fp32, build_csr)
        self.op64 = self.op32 = self.csr = None
        self.assembled = False
        if assemble:
            self.assemble_primitives()
            self.finish_assembly()

    def assemble_primitives(self):
        """(Re)assemble the FP64 primitives of this instance alone
        (mpet_assemble_primitives), before any Robin term is added."""
        check(_lib.lib().mpet_assemble_primitives(
            C.byref(self.mesh_desc), C.byref(self.pat), ptr(self.Kcell), self.A, self.mu,
            C.byref(self.prim), _lib.stream_ptr()))

    def finish_assembly(self):
        """After the primitives are assembled (here or by assemble_batch):
        Robin terms, structured copies, scalar CSR."""
        robin, structured, need_fp32, build_csr = self._rest
        self._add_robin(robin)
        # structured operator copies (a batch packs its own from the primitives)
        self.op64 = self._structured(0) if structured else None
        self.op32 = self._structured(1) if structured and need_fp32 else None
        self.csr = self.expand_csr(0) if build_csr else None
        self.assembled = True

    # -- assembly --------------------------------------------------------------
    def _alloc_primitives(self):
        dev = self.device
        f64 = torch.float64
        nz = self.nnz
        self.prim_t = {
            "Kuu": torch.empty(9 * nz["22"], dtype=f64, device=dev),
            "Lap": torch.empty(nz["22"], dtype=f64, device=dev),
            "Dv": torch.empty(3 * nz["21"], dtype=f64, device=dev),
            "Dt": torch.empty(3 * nz["12"], dtype=f64, device=dev),
            "M1": torch.empty(nz["11"], dtype=f64, device=dev),
            "St": torch.empty(self.A * nz["11"], dtype=f64, device=dev),
        }
        t = self.prim_t
        self.prim = _lib.Primitives(ptr(t["Kuu"]), ptr(t["Lap"]), ptr(t["Dv"]), ptr(t["Dt"]),
                                    ptr(t["M1"]), ptr(t["St"]))
        self.primK = self.prim

    def _add_robin(self, robin):
        """Robin terms of lhs(L2_i) (:324, added to A at :439-441): A gets
        -dt theta int_{ds_i(2)} beta_i p_i w_i, i.e. St_i += int beta_i psi psi
        (St_i is scaled by -dt theta in A).  The preconditioner prec (:279-283)
        and the up_-linear right-hand side keep the Robin-free St ("StK")."""
        t = self.prim_t
        t["StK"] = t["St"]
        if not robin:
            return
        from .facets import robin_matrix
        t["StK"] = t["St"].clone()
        n11 = self.nnz["11"]
        for i, fs, beta, time in robin:
            robin_matrix(fs, beta, time, self.pat_dev["11"], t["St"][i * n11:(i + 1) * n11])
        self.primK = _lib.Primitives(ptr(t["Kuu"]), ptr(t["Lap"]), ptr(t["Dv"]), ptr(t["Dt"]),
                                     ptr(t["M1"]), ptr(t["StK"]))

    def st_row(self, f):
        """Row of St (network) carrying the stiffness of pressure field f."""
        return f - 1 if f >= 1 else None

    def _structured(self, dtype):
        """SoA copy of the primitives in float64 (dtype 0) or float32 (1)."""
        tdt = torch.float64 if dtype == 0 else torch.float32
        L = _lib.lib()
        s = _lib.stream_ptr()
        out = {}
        for name, ncomp in (("Kuu", 9), ("Dv", 3), ("Dt", 3), ("M1", 1)):
            src = self.prim_t[name]
            dst = torch.empty_like(src, dtype=tdt)
            check(L.mpet_soa_convert(dtype, ptr(src), ncomp, src.numel() // ncomp, ptr(dst), s))
            out[name] = dst
        out["St"] = self.prim_t["St"].to(tdt)
        return out

    lanes = {0: (16, 16), 1: (16, 16)}     # lanes per (P2, P1) node row, per dtype
    # row tiles per XCD run, per dtype (0: off); runs of 4 make the FP32
    # apply 4.5% faster at 48^3 (0.303 vs 0.317 ms), flat from 2 to 128
    xcd_chunk = {0: 4, 1: 4}

    def op_desc(self, dtype, masked=True, lanes=None, xcd_chunk=None):
        vals = self.op64 if dtype == 0 else self.op32
        coef = _lib.make_coef(self.Cm, self.Cs, self.mu)
        lu, lp = lanes or self.lanes[dtype]
        xc = self.xcd_chunk[dtype] if xcd_chunk is None else xcd_chunk
        return _lib.OpDesc(self.nP2, self.nP1, self.A, self.pat, dtype, ptr(vals["Kuu"]),
                           ptr(vals["Dv"]), ptr(vals["Dt"]), ptr(vals["M1"]), ptr(vals["St"]),
                           coef, ptr(self.dmask) if masked else None, lu, lp, xc)

    # lanes per (P2, P1) node row and XCD run length of the cell-wise
    # operator's row kernel
    cell_lanes = (8, 8)
    cell_xcd_chunk = 4

    CELLS_PER_BLOCK = 256          # cells of one block of the cell kernel (CW_NT)

    def cell_op(self):
        """Arrays of the cell-wise FP32 operator, built once: per-cell
        geometry (FP32 gradients of the barycentric coordinates and |T|),
        the cell -> node tables in structure-of-arrays layout, and the
        block partials: blocks of CELLS_PER_BLOCK consecutive cells write
        one partial per (node, block), at its entry in the list of all
        partials sorted by node then block, summing the node's local slots
        of the block in slot order."""
        if getattr(self, "_cw", None) is None:
            dev = self.device
            nC = int(self.cells.shape[0])
            B = self.CELLS_PER_BLOCK
            geom = torch.empty(10 * nC, dtype=torch.float32, device=dev)
            check(_lib.lib().mpet_cell_geometry(C.byref(self.mesh_desc), ptr(geom),
                                                _lib.stream_ptr()))
            self._cw = {"nC": nC, "geom": geom,
                        "p2": block_partial_tables(self.cell_p2, self.nP2, B),
                        "p1": block_partial_tables(self.cells, self.nP1, B)}
        return self._cw

    def cell_work_floats(self):
        """Work values of the cell-wise operator (its partials), either precision."""
        cw = self.cell_op()
        return 3 * cw["p2"]["npart"] + cw["p1"]["npart"]

    def cell_geom64(self):
        """FP64 per-cell geometry of the FP64 cell-wise operator, built once."""
        cw = self.cell_op()
        if "geom64" not in cw:
            g = torch.empty(10 * cw["nC"], dtype=torch.float64, device=self.device)
            check(_lib.lib().mpet_cell_geometry64(C.byref(self.mesh_desc), ptr(g),
                                                  _lib.stream_ptr()))
            cw["geom64"] = g
        return cw["geom64"]

    def cell_op_desc(self, masked=True, lanes=None, xcd_chunk=None, fp64=False):
        """MpetCellOpDesc of the cell-wise operator: FP32 (the P1 x P1 block
        from the FP32 structured copy) or FP64 (FP64 geometry, the P1 x P1
        block from the FP64 primitives)."""
        if fp64:
            v, geom = self.op64, self.cell_geom64()
        else:
            if self.op32 is None:
                raise ValueError("the cell-wise operator needs the FP32 structured copy")
            v, geom = self.op32, self.cell_op()["geom"]
        cw = self.cell_op()
        t2, t1 = cw["p2"], cw["p1"]
        lu, lp = lanes or self.cell_lanes
        xc = self.cell_xcd_chunk if xcd_chunk is None else xcd_chunk
        p11 = self.pat_dev["11"]
        return _lib.CellOpDesc(cw["nC"], self.nP2, self.nP1, self.A, ptr(t2["node"]),
                               ptr(t1["node"]), ptr(geom), self.CELLS_PER_BLOCK,
                               t2["npart"], t1["npart"], ptr(t2["bq"]), ptr(t2["lrp"]),
                               ptr(t2["pe"]), ptr(t2["lsl"]), ptr(t1["bq"]), ptr(t1["lrp"]),
                               ptr(t1["pe"]), ptr(t1["lsl"]), ptr(t2["rp"]), ptr(t1["rp"]),
                               ptr(p11[0]), ptr(p11[1]), self.nnz["11"], ptr(v["M1"]),
                               ptr(v["St"]), _lib.make_coef(self.Cm, self.Cs, self.mu),
                               ptr(self.dmask) if masked else None, lu, lp, xc, int(fp64))

    def cell_apply(self, x32, b=None, masked=True, lanes=None):
        """y = A x (or b - A x) through the cell-wise operator, in the
        precision of x (FP32 or FP64)."""
        y = torch.empty_like(x32)
        fp64 = x32.dtype == torch.float64
        work = torch.empty(self.cell_work_floats(), dtype=x32.dtype, device=self.device)
        d = self.cell_op_desc(masked, lanes, fp64=fp64)
        check(_lib.lib().mpet_cell_op_apply(C.byref(d), ptr(x32), ptr(b), ptr(y), ptr(work),
                                            _lib.stream_ptr()))
        return y

    def row_lengths(self, which):
        rp = {k: np.diff(v[0]) for k, v in self.pat_host.items()}
        F = self.F
        lu = (3 * rp["22"] + (rp["21"] if which == 0 else 0))
        lens_u = np.repeat(lu, 3)
        lp = np.empty((self.nP1, F), dtype=np.int64)
        if which == 0:
            lp[:, :] = F * rp["11"][:, None]
            lp[:, 0] += 3 * rp["12"]
        else:
            lp[:, :] = rp["11"][:, None]
        return np.concatenate([lens_u, lp.reshape(-1)])

    def expand_csr(self, which):
        """Scalar CSR (rowptr int64, cols int32, vals f64) of the
        un-eliminated A (which=0) or P (which=1)."""
        lens = self.row_lengths(which)
        rowptr = np.zeros(self.n + 1, dtype=np.int64)
        np.cumsum(lens, out=rowptr[1:])
        nnz = int(rowptr[-1])
        if nnz >= 2**31:
            raise ValueError("scalar CSR too large for int32 column indices")
        dev = self.device
        rp = _dev(rowptr, torch.int64, dev)
        cols = torch.empty(nnz, dtype=torch.int32, device=dev)
        vals = torch.empty(nnz, dtype=torch.float64, device=dev)
        if which == 0:
            coef = _lib.make_coef(self.Cm, self.Cs, self.mu)
        else:
            coef = _lib.make_coef(np.diag(self.Pd), self.Ps, self.mu)
        prim = self.prim if which == 0 else self.primK
        check(_lib.lib().mpet_expand_csr(C.byref(self.pat), C.byref(prim), self.nP2,
                                         self.nP1, self.A, C.byref(coef), which, ptr(rp),
                                         ptr(cols), ptr(vals), _lib.stream_ptr()))
        return rp, cols, vals

    # -- Dirichlet data ----------------------------------------------------------
    def set_dirichlet(self, dofs):
        m = torch.zeros(self.n, dtype=torch.uint8)
        if len(dofs):
            m[torch.as_tensor(np.asarray(dofs, dtype=np.int64))] = 1
        self.dmask.copy_(m.to(self.device))

    # -- operator applications ----------------------------------------------------
    def apply(self, x, b=None, dtype=0, masked=True):
        y = torch.empty_like(x)
        d = self.op_desc(dtype, masked)
        check(_lib.lib().mpet_op_apply(C.byref(d), ptr(x), ptr(b), ptr(y), _lib.stream_ptr()))
        return y

    def csr_apply(self, x, b=None, lpr=16, masked=True, csr=None):
        rp, ci, vals = csr if csr is not None else self.csr
        y = torch.empty_like(x)
        check(_lib.lib().mpet_csr_spmv(0, self.n, ptr(rp), ptr(ci), ptr(vals), ptr(x), ptr(b),
                                       ptr(self.dmask) if masked else None, ptr(y), lpr,
                                       _lib.stream_ptr()))
        return y

    def cell_source(self, test_degree, Q, fval, ncomp, scale, dof_base, dof_stride, dof_off, b):
        """b += scale * int f_h . v over all cells (f_h: per-cell P_k nodal
        values fval (nC*nk*ncomp), Q: reference weights (ntest, nk))."""
        test = self.cell_p2 if test_degree == 2 else self.cells
        ntest = 10 if test_degree == 2 else 4
        nk = Q.shape[1]
        check(_lib.lib().mpet_cell_source(C.byref(self.mesh_desc), ptr(test), ntest, ptr(Q), nk,
                                          ptr(fval), ncomp, float(scale), int(dof_base),
                                          int(dof_stride), int(dof_off), ptr(b),
                                          _lib.stream_ptr()))

    def rhs_pressure(self, up_prev, b):
        coef = _lib.make_coef(self.Bm, self.Bs, self.mu)
        check(_lib.lib().mpet_rhs_pressure(C.byref(self.pat), C.byref(self.primK), self.nP2,
                                           self.nP1, self.A, C.byref(coef), ptr(up_prev),
                                           ptr(b), _lib.stream_ptr()))
        return b


class TwoFieldSystem(DeviceSystem):
    """Assembled two-field MPET system (MPETSolver, mpet/mpetsolver.py:196-201,
    fields u, p_1..p_A) of one instance on one GPU, as scalar FP64 CSR
    matrices built on the device from the same primitives as the
    total-pressure system (`mpet_tf_expand_csr`)."""

    two_field = True

    def __init__(self, space: MixedSpace, params, dt, theta, K_cells=None, device="cuda",
                 robin=None):
        if space.total_pressure:
            raise ValueError("TwoFieldSystem needs a space without total pressure")
        super().__init__(space, params, dt, theta, K_cells=K_cells, device=device,
                         build_csr=False, need_fp32=False, structured=False, robin=robin)
        A = space.A
        self.F = A
        c, S = params["c"], params["S"]
        self.alpha = np.array([float(a) for a in params["alpha"]])
        self.Cm, self.Cs = forms.two_field_matrix_coefficients(c, S, dt, theta)
        self.Bm, self.Bs = forms.two_field_rhs_coefficients(c, S, dt, theta)
        self.Pd, self.Ps = forms.two_field_preconditioner_coefficients(c, S, dt, theta)
        self.csr = self.expand_csr(0)
        self.Bcsr = self.expand_csr(2)

    def st_row(self, f):
        return f

    def op_desc(self, dtype, masked=True, lanes=None):
        raise NotImplementedError("the two-field system is held as scalar CSR only")

    def row_lengths(self, which):
        rp = {k: np.diff(v[0]) for k, v in self.pat_host.items()}
        F = self.F
        if which == 0:
            lu = 3 * rp["22"] + F * rp["21"]
        elif which == 1:
            lu = 3 * rp["22"]
        else:
            lu = np.zeros(self.nP2, dtype=np.int64)
        lens_u = np.repeat(lu, 3)
        lp = np.empty((self.nP1, F), dtype=np.int64)
        lp[:, :] = (rp["11"] if which == 1 else 3 * rp["12"] + F * rp["11"])[:, None]
        return np.concatenate([lens_u, lp.reshape(-1)])

    def expand_csr(self, which):
        """Scalar CSR (rowptr int64, cols int32, vals f64): which = 0 the
        matrix of lhs(F) with Robin, 1 the preconditioner prec, 2 the matrix
        of the up_-linear part of rhs(F)."""
        lens = self.row_lengths(which)
        rowptr = np.zeros(self.n + 1, dtype=np.int64)
        np.cumsum(lens, out=rowptr[1:])
        nnz = int(rowptr[-1])
        if nnz >= 2**31:
            raise ValueError("scalar CSR too large for int32 column indices")
        dev = self.device
        rp = _dev(rowptr, torch.int64, dev)
        cols = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
        vals = torch.empty(max(nnz, 1), dtype=torch.float64, device=dev)
        if which == 0:
            Cm, st, prim = self.Cm, self.Cs, self.prim
        elif which == 1:
            Cm, st, prim = np.diag(self.Pd), self.Ps, self.primK
        else:
            Cm, st, prim = self.Bm, self.Bs, self.primK
        coef = _lib.make_two_field_coef(Cm, st, self.alpha, self.mu, self.lmbda)
        check(_lib.lib().mpet_tf_expand_csr(C.byref(self.pat), C.byref(prim), self.nP2,
                                            self.nP1, self.A, C.byref(coef), which, ptr(rp),
                                            ptr(cols), ptr(vals), _lib.stream_ptr()))
        return rp, cols, vals

    def rhs_pressure(self, up_prev, b):
        """b = (matrix of the up_-linear part of rhs(F)) up_ — `assemble(L)`
        (:356, :433); the u rows are 0."""
        rp, ci, vals = self.Bcsr
        check(_lib.lib().mpet_csr_spmv(0, self.n, ptr(rp), ptr(ci), ptr(vals), ptr(up_prev),
                                       None, None, ptr(b), 16, _lib.stream_ptr()))
        return b
