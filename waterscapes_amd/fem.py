"""Function spaces, DOF maps and coefficient objects of the MPET mixed space.

The mixed space is [P2]^3 x P1^(A+1) with field order u, p_0 (total
pressure), p_1..p_A (mpet/mpettotalpressuresolver.py:175-204).  Canonical
global numbering used by every backend and by the FP64 oracle:

  * P2 nodes  = vertices + edge midpoints, renumbered in lexicographic
    (z, y, x) coordinate order (on a box this is the (2n+1)^3 lattice);
  * P1 nodes  = mesh vertices (already lexicographic on a box);
  * u DOF     = 3 * a + j                          (a P2 node, j component);
  * p_f DOF   = 3 * nP2 + (A + 1) * c + f          (c P1 node, f = 0..A).

The two-field space of MPETSolver (u, p_1..p_A; mpet/mpetsolver.py:100-131)
uses the same numbering with F = A pressure fields per node in place of A+1.
Components of u and the pressures of one node are adjacent in memory, so
3x3 blocks and the pressure vector of a node are read with one contiguous
load.

``Constant``, ``Expression``, ``Function``, ``interpolate``, ``assign`` and
``errornorm`` are the DOLFIN look-alikes the reference API exposes.
"""
from __future__ import annotations

import math
import re

import numpy as np

from . import refel
from .mesh import Mesh

# ---------------------------------------------------------------------------
# DOF map
# ---------------------------------------------------------------------------


class DofMap:
    """Cell -> P2 / P1 node maps, node coordinates and node patterns."""

    def __init__(self, mesh: Mesh, lexicographic=True):
        self.mesh = mesh
        nV = mesh.num_vertices()
        nE = mesh.num_edges()
        x = mesh.coordinates()
        cells = mesh.cells()
        # unordered P2 nodes: vertices then edges
        c2 = np.concatenate([cells, nV + mesh.cell_edges], axis=1)        # nC,10
        x2 = np.concatenate([x, 0.5 * (x[mesh.edges[:, 0]] + x[mesh.edges[:, 1]])])
        if lexicographic:
            order = np.lexsort((x2[:, 0], x2[:, 1], x2[:, 2]))           # new -> old
            inv = np.empty_like(order)
            inv[order] = np.arange(order.size)
            c2 = inv[c2]
            x2 = x2[order]
        self.nP2 = nV + nE
        self.nP1 = nV
        self.cell_p2 = np.ascontiguousarray(c2.astype(np.int32))
        self.cell_p1 = np.ascontiguousarray(cells.astype(np.int32))
        self.x2 = np.ascontiguousarray(x2)
        self.x1 = x
        self._patterns = {}

    # node patterns (CSR, sorted columns) from cell connectivity
    def pattern(self, kind):
        """kind in {'22','21','12','11'}: rows / cols = P2 or P1 nodes."""
        if kind in self._patterns:
            return self._patterns[kind]
        rmap = self.cell_p2 if kind[0] == "2" else self.cell_p1
        cmap = self.cell_p2 if kind[1] == "2" else self.cell_p1
        nrow = self.nP2 if kind[0] == "2" else self.nP1
        ncol = self.nP2 if kind[1] == "2" else self.nP1
        rowptr, cols = build_pattern(rmap, cmap, nrow, ncol)
        self._patterns[kind] = (rowptr, cols)
        return rowptr, cols


def build_pattern(rmap, cmap, nrow, ncol):
    """CSR (rowptr int64, cols int32) of all (r, c) with r, c in one cell
    (native builder in csrc/setup_host.cpp; numpy fallback for hosts that
    have not built the library, used only by host-side tests)."""
    try:
        from . import _lib
        L = _lib.lib()
    except RuntimeError:
        return _build_pattern_np(rmap, cmap, nrow, ncol)
    import ctypes as C
    rmap = np.ascontiguousarray(rmap, dtype=np.int32)
    cmap = np.ascontiguousarray(cmap, dtype=np.int32)
    rowptr = np.zeros(nrow + 1, dtype=np.int64)
    fn = L.mpet_build_pattern
    fn.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                   C.c_void_p]
    fn.restype = C.c_int
    _lib.check(fn(rmap.shape[0], rmap.shape[1], rmap.ctypes.data, cmap.shape[1],
                  cmap.ctypes.data, nrow, rowptr.ctypes.data, None))
    cols = np.empty(int(rowptr[-1]), dtype=np.int32)
    _lib.check(fn(rmap.shape[0], rmap.shape[1], rmap.ctypes.data, cmap.shape[1],
                  cmap.ctypes.data, nrow, rowptr.ctypes.data, cols.ctypes.data))
    return rowptr, cols


def _build_pattern_np(rmap, cmap, nrow, ncol):
    nC = rmap.shape[0]
    nr, nc = rmap.shape[1], cmap.shape[1]
    keys = (rmap.astype(np.int64)[:, :, None] * ncol
            + cmap.astype(np.int64)[:, None, :]).reshape(-1)
    keys = np.unique(keys)
    rows = keys // ncol
    cols = (keys % ncol).astype(np.int32)
    rowptr = np.zeros(nrow + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=nrow), out=rowptr[1:])
    del nC, nr, nc
    return rowptr, cols


def cell_colouring(mesh: Mesh):
    """Colour cells so that no two cells of one colour share a vertex (hence
    no P2 or P1 node).  Returns (perm, offsets): cells of colour k are
    perm[offsets[k]:offsets[k+1]].  Box meshes use cube parity (8 classes)
    times the tetrahedron index inside the cube (6); others a greedy colouring
    in cell order."""
    nC = mesh.num_cells()
    if mesh.structured is not None:
        nx, ny, nz = mesh.structured
        cube = np.arange(nC) // 6
        tet = np.arange(nC) % 6
        ix = cube % nx
        iy = (cube // nx) % ny
        iz = cube // (nx * ny)
        col = ((ix % 2) + 2 * (iy % 2) + 4 * (iz % 2)) * 6 + tet
    else:
        col = _greedy_colour(mesh)
    perm = np.argsort(col, kind="stable").astype(np.int32)
    ncol = int(col.max()) + 1
    offsets = np.zeros(ncol + 1, dtype=np.int64)
    np.cumsum(np.bincount(col, minlength=ncol), out=offsets[1:])
    return perm, offsets


def _greedy_colour(mesh):
    cells = mesh.cells()
    nV = mesh.num_vertices()
    # vertex -> cells
    flat = cells.reshape(-1)
    order = np.argsort(flat, kind="stable")
    vptr = np.zeros(nV + 1, dtype=np.int64)
    np.cumsum(np.bincount(flat, minlength=nV), out=vptr[1:])
    vcells = order // 4
    col = np.full(mesh.num_cells(), -1, dtype=np.int64)
    for e in range(mesh.num_cells()):
        used = set()
        for v in cells[e]:
            used.update(col[vcells[vptr[v]:vptr[v + 1]]].tolist())
        k = 0
        while k in used:
            k += 1
        col[e] = k
    return col


# ---------------------------------------------------------------------------
# coefficients
# ---------------------------------------------------------------------------


class Constant:
    """Spatially constant coefficient; ``time`` is a Constant in the
    reference (mpet/mpetproblem.py:109-111)."""

    def __init__(self, value):
        self._v = np.array(value, dtype=np.float64)

    def assign(self, value):
        self._v = np.array(float(value) if np.ndim(value) == 0 else value,
                           dtype=np.float64)

    def values(self):
        return np.atleast_1d(self._v).copy()

    def value_shape(self):
        return self._v.shape

    def __float__(self):
        return float(self._v)

    def __repr__(self):
        return f"Constant({self._v.tolist()})"

    # evaluation protocol shared with Expression
    degree = 0

    def is_zero(self):
        return not np.any(self._v)

    def depends_on_time(self):
        return False

    def eval(self, X, t=None):
        n = X.shape[0]
        v = self._v
        if v.ndim == 0:
            return np.full(n, float(v))
        return np.broadcast_to(v, (n,) + v.shape).copy()


_CFUNCS = {"pow": "Pow", "fabs": "Abs", "abs": "Abs"}


def _torch_functions():
    """The functions sympy prints for C expressions, as torch operations."""
    import torch
    return {"sin": torch.sin, "cos": torch.cos, "tan": torch.tan, "asin": torch.asin,
            "acos": torch.acos, "atan": torch.atan, "atan2": torch.atan2, "sinh": torch.sinh,
            "cosh": torch.cosh, "tanh": torch.tanh, "exp": torch.exp, "log": torch.log,
            "sqrt": torch.sqrt, "Abs": torch.abs, "floor": torch.floor,
            "ceiling": torch.ceil, "Min": torch.minimum, "Max": torch.maximum,
            "sign": torch.sign, "erf": torch.erf}


def _cpp_to_sympy(src):
    s = re.sub(r"x\[(\d)\]", r"x\1", src)
    s = s.replace("M_PI", "pi")
    for c, p in _CFUNCS.items():
        s = re.sub(r"\b%s\(" % c, p + "(", s)
    return s


class Expression:
    """``Expression(code, degree=k, t=time, **params)``.

    ``code`` is a C++ expression string (as produced by sympy.printing.ccode
    in the reference tests, e.g. "sin(2*pi*x[0])*t"), a tuple of them for a
    vector / nested tuples for a tensor, or a Python callable ``f(x, t)`` with
    x an (n, 3) array returning (n,) or (n, vdim).  Like DOLFIN, a source with
    ``degree=k`` is interpolated into P_k on each cell before integration;
    boundary values and initial conditions are nodal evaluations."""

    def __init__(self, code, degree=2, element=None, **params):
        import sympy
        self.degree = int(degree)
        self.params = dict(params)
        self._callable = None
        if callable(code):
            self._callable = code
            self.shape = None
            return
        arr = np.array(code, dtype=object)
        self.shape = arr.shape
        xs = sympy.symbols("x0 x1 x2")
        names = {"x0": xs[0], "x1": xs[1], "x2": xs[2], "pi": sympy.pi}
        self._syms = {}
        for k in self.params:
            self._syms[k] = sympy.Symbol(k)
            names[k] = self._syms[k]
        exprs = [sympy.sympify(_cpp_to_sympy(str(c)), locals=names)
                 for c in arr.reshape(-1)]
        self._exprs = exprs
        args = list(xs) + [self._syms[k] for k in self.params]
        self._np = [sympy.lambdify(args, e, modules="numpy") for e in exprs]
        free = set()
        for e in exprs:
            free |= {str(s) for s in e.free_symbols}
        self._time_dep = any(k in free and isinstance(v, Constant)
                             for k, v in self.params.items())
        self._zero = all(e == 0 for e in exprs)

    def is_zero(self):
        return self._callable is None and self._zero

    def depends_on_time(self):
        return self._callable is not None or self._time_dep

    def value_shape(self):
        return self.shape

    def _param_values(self, t):
        vals = []
        for k, v in self.params.items():
            if isinstance(v, Constant):
                vals.append(float(v) if (t is None or k != "t") else float(t))
            else:
                vals.append(float(v))
        return vals

    def eval_device(self, X, t=None):
        """eval() on a device tensor X (n, 3) float64, computed with torch
        operations where X lives (the per-step source update stays on the
        GPU).  A Python-callable Expression is evaluated on the host and
        copied, since its code is opaque."""
        import torch
        n = X.shape[0]
        if self._callable is not None:
            v = self.eval(X.detach().cpu().numpy(), t)
            return torch.as_tensor(np.ascontiguousarray(v, dtype=np.float64)).to(X.device)
        if getattr(self, "_torch", None) is None:
            import sympy
            xs = sympy.symbols("x0 x1 x2")
            args = list(xs) + [self._syms[k] for k in self.params]
            self._torch = [sympy.lambdify(args, e, modules=[_torch_functions(), "math"])
                           for e in self._exprs]
        # parameters as 0-d device tensors: functions of t alone then run in
        # torch too (a float argument is not accepted by torch.sin etc.)
        pv = [torch.tensor(v, dtype=torch.float64, device=X.device)
              for v in self._param_values(t)]
        out = []
        for f in self._torch:
            v = f(X[:, 0], X[:, 1], X[:, 2], *pv)
            if not isinstance(v, torch.Tensor):
                v = torch.full((n,), float(v), dtype=torch.float64, device=X.device)
            out.append(v.to(torch.float64).expand(n))
        if self.shape == ():
            return out[0].contiguous()
        return torch.stack(out, dim=1).reshape((n,) + tuple(self.shape))

    def eval(self, X, t=None):
        """Values at points X (n, 3) -> (n,) or (n,)+shape (float64).
        ``t`` overrides the value of a Constant parameter named t."""
        X = np.asarray(X, dtype=np.float64)
        n = X.shape[0]
        if self._callable is not None:
            tt = t
            if tt is None and "t" in self.params:
                tt = float(self.params["t"])
            return np.asarray(self._callable(X, tt), dtype=np.float64)
        pv = self._param_values(t)
        out = [np.broadcast_to(np.asarray(f(X[:, 0], X[:, 1], X[:, 2], *pv),
                                          dtype=np.float64), (n,))
               for f in self._np]
        if self.shape == ():
            return out[0].copy()
        return np.stack(out, axis=1).reshape((n,) + self.shape)


class FacetNormalTraction:
    """s = sigma . n for a tensor-valued Expression sigma, or sigma n for a
    scalar one; produced by ``expression * FacetNormal(mesh)`` (reference
    test_convergence_totalpressuresolver.py:135-138, test_donut.py:45)."""

    def __init__(self, sigma):
        self.sigma = sigma
        self.degree = sigma.degree

    def is_zero(self):
        return self.sigma.is_zero()

    def depends_on_time(self):
        return self.sigma.depends_on_time()

    def eval_normal(self, X, normals, t=None):
        s = np.asarray(self.sigma.eval(X, t))
        if s.ndim == 1:                                # scalar * n (test_donut.py:45)
            return s[:, None] * normals
        return np.einsum("nij,nj->ni", s, normals)     # n,3,3

    def eval_normal_device(self, X, normals, t=None):
        """eval_normal on device tensors (sigma evaluated with torch)."""
        import torch
        s = self.sigma.eval_device(X, t)
        if s.dim() == 1:
            return s[:, None] * normals
        return torch.einsum("nij,nj->ni", s, normals)


class FacetNormal:
    def __init__(self, mesh):
        self.mesh = mesh

    def __rmul__(self, other):
        if isinstance(other, Expression):
            return FacetNormalTraction(other)
        raise TypeError("only Expression * FacetNormal is supported")


Expression.__mul__ = lambda self, other: (FacetNormalTraction(self)
                                          if isinstance(other, FacetNormal)
                                          else NotImplemented)

# ---------------------------------------------------------------------------
# spaces and functions
# ---------------------------------------------------------------------------


class MixedSpace:
    """[P2]^3 x P1^F on a mesh, canonical numbering (module docstring):
    F = A + 1 fields u, p_0, p_1..p_A of the total-pressure formulation, or
    (total_pressure=False) F = A fields u, p_1..p_A of the two-field
    MPETSolver (mpet/mpetsolver.py:100-131).

    Nullspace multipliers (the R-space components the reference appends to
    the mixed element, mpettotalpressuresolver.py:182-202, mpetsolver.py:
    107-128): u_nullspace adds one 6-vector (rigid motions), p_fields lists
    the pressure fields (field numbering of the layout: the reference's
    p[k]) that get one scalar multiplier each.  Their coefficients follow the
    ndofs field DOFs; dim() counts both."""

    def __init__(self, mesh, A, dofmap=None, total_pressure=True, u_nullspace=False,
                 p_fields=()):
        self.mesh = mesh
        self.A = int(A)
        self.total_pressure = bool(total_pressure)
        self.F = self.A + 1 if self.total_pressure else self.A
        self.dofmap = dofmap if dofmap is not None else DofMap(mesh)
        self.nP2 = self.dofmap.nP2
        self.nP1 = self.dofmap.nP1
        self.nu = 3 * self.nP2
        self.np = self.F * self.nP1
        self.ndofs = self.nu + self.np
        self.u_nullspace = bool(u_nullspace)
        self.p_fields = tuple(int(f) for f in p_fields)
        if any(f < 0 or f >= self.F for f in self.p_fields):
            raise ValueError("p_fields: pressure field index out of range")
        self.mult_groups = ([6] if self.u_nullspace else []) + [1] * len(self.p_fields)
        self.nmult = sum(self.mult_groups)

    def dim(self):
        return self.ndofs + self.nmult

    def sub(self, i):
        return SubSpace(self, i)

    def num_sub_spaces(self):
        return self.F + 1 + len(self.mult_groups)

    def network_sub(self, i):
        """Sub-space index of network pressure p_{i+1} (i 0-based): i + 2 with
        a total pressure, i + 1 without."""
        return i + 1 + (self.F - self.A)

    def field_index(self, i):
        """Global DOF indices of field i (0 = u (nP2*3, node-major); then the
        pressure fields in order: p_0, p_1..p_A, or p_1..p_A)."""
        if i == 0:
            return np.arange(self.nu)
        if i > self.F:
            j = i - self.F - 1
            off = self.ndofs + sum(self.mult_groups[:j])
            return np.arange(off, off + self.mult_groups[j])
        return self.nu + np.arange(self.nP1) * self.F + (i - 1)


class SubSpace:
    def __init__(self, parent, i):
        self.parent = parent
        self.i = i

    def collapse(self):
        return CollapsedSpace(self.parent, self.i)

    def dim(self):
        return len(self.parent.field_index(self.i))


class CollapsedSpace:
    """Standalone P2-vector (i == 0) or P1 space of a mixed-space field."""

    def __init__(self, parent, i):
        self.parent = parent
        self.i = i
        self.mesh = parent.mesh

    def dim(self):
        return self.parent.nu if self.i == 0 else self.parent.nP1

    def node_coordinates(self):
        dm = self.parent.dofmap
        return dm.x2 if self.i == 0 else dm.x1


class _Vector:
    def __init__(self, fn):
        self._f = fn

    def get_local(self):
        return self._f._array_np().copy()

    def set_local(self, v):
        self._f._set(np.asarray(v, dtype=np.float64))

    def size(self):
        return self._f._size()

    def __setitem__(self, idx, v):
        a = self._f._array_np().copy()
        a[idx] = v
        self._f._set(a)

    def __getitem__(self, idx):
        return self._f._array_np()[idx]

    def norm(self, kind="l2"):
        a = self._f._array_np()
        return float(np.linalg.norm(a) if kind == "l2" else np.abs(a).max())


class Function:
    """Coefficient vector on a MixedSpace or CollapsedSpace.  The storage is
    a float64 numpy array or torch tensor (device backends keep it on the
    GPU; ``vector().get_local()`` copies to the host)."""

    def __init__(self, space, data=None):
        self._space = space
        if data is None:
            data = np.zeros(space.dim())
        self.data = data

    def function_space(self):
        return self._space

    def vector(self):
        return _Vector(self)

    def _array_np(self):
        d = self.data
        if isinstance(d, np.ndarray):
            return d
        return d.detach().cpu().numpy()

    def _size(self):
        return int(self.data.shape[0])

    def _set(self, arr):
        if isinstance(self.data, np.ndarray):
            self.data[:] = arr
        else:
            import torch
            self.data.copy_(torch.as_tensor(arr, dtype=self.data.dtype))

    def assign(self, other):
        src = other.data if isinstance(other, Function) else other
        if isinstance(self.data, np.ndarray):
            self.data[:] = src if isinstance(src, np.ndarray) else src.cpu().numpy()
        else:
            import torch
            self.data.copy_(torch.as_tensor(src).to(self.data.device,
                                                     self.data.dtype))

    def sub(self, i):
        return FieldView(self, i)

    def split(self, deepcopy=False):
        sp = self._space
        if not isinstance(sp, MixedSpace):
            return (self,)
        return tuple(FieldView(self, i) for i in range(sp.num_sub_spaces()))

    def copy(self, deepcopy=True):
        d = self.data.copy() if isinstance(self.data, np.ndarray) else self.data.clone()
        return Function(self._space, d)

    # values at nodes, per field (host arrays)
    def field_values(self, i):
        sp = self._space
        a = self._array_np()
        if not isinstance(sp, MixedSpace):
            return a.reshape(-1, 3) if sp.i == 0 else a
        if i == 0:
            return a[: sp.nu].reshape(sp.nP2, 3)
        if i > sp.F:                                  # nullspace multipliers
            return a[sp.field_index(i)]
        return a[sp.nu: sp.ndofs].reshape(sp.nP1, sp.F)[:, i - 1]


class FieldView:
    """up.split()[i] / up.sub(i): a view of one field of a mixed Function."""

    def __init__(self, parent, i):
        self.parent = parent
        self.i = i

    def function_space(self):
        return self.parent.function_space().sub(self.i)

    def values(self):
        return self.parent.field_values(self.i)

    def vector(self):
        return _FieldVector(self)


class _FieldVector:
    def __init__(self, fv):
        self.fv = fv

    def get_local(self):
        return np.asarray(self.fv.values()).reshape(-1).copy()


def interpolate(expr, space):
    """Nodal interpolation of an Expression / Constant into a collapsed
    P2-vector or P1 space (DOLFIN interpolate on Lagrange spaces)."""
    X = space.node_coordinates()
    v = expr.eval(X)
    if space.i == 0:
        v = np.broadcast_to(v, (X.shape[0], 3)).reshape(-1)
    return Function(space, np.ascontiguousarray(v, dtype=np.float64))


def assign(target, source):
    """assign(up_.sub(i), interpolate(...)) — copy a collapsed-space Function
    into field i of a mixed Function."""
    if not isinstance(target, FieldView):
        target.assign(source)
        return
    mixed = target.parent
    sp = mixed.function_space()
    vals = source._array_np() if isinstance(source, Function) else np.asarray(source)
    a = mixed._array_np().copy()
    a[sp.field_index(target.i)] = vals
    mixed._set(a)


def errornorm(exact, approx, norm_type="L2", degree_rise=3, t=None):
    """||exact - approx|| in L2 or H1 (full norm) over the mesh, computed with a
    degree-8 quadrature per cell; ``approx`` is a FieldView of a mixed
    Function (P2 vector for field 0, P1 otherwise)."""
    sp = approx.parent.function_space()
    mesh = sp.mesh
    dm = sp.dofmap
    qx, qw = tet_quadrature(8)
    x = mesh.coordinates()
    c = mesh.cells()
    v0 = x[c[:, 0]]
    J = np.stack([x[c[:, 1]] - v0, x[c[:, 2]] - v0, x[c[:, 3]] - v0], axis=2)
    det = np.abs(np.linalg.det(J))
    G = np.linalg.inv(J)                    # rows = grad lambda_1..3
    grads = np.concatenate([-G.sum(axis=1, keepdims=True), G], axis=1)  # nC,4,3
    lam = np.concatenate([1 - qx.sum(1, keepdims=True), qx], axis=1)  # nq,4
    X = v0[:, None, :] + np.einsum("cij,qj->cqi", J, qx)               # nC,nq,3
    nC, nq = X.shape[:2]
    if approx.i == 0:
        basis = p2_values(lam)                                          # nq,10
        dbasis = p2_grads(lam, grads)                                   # nC,nq,10,3
        coef = approx.values()[dm.cell_p2]                              # nC,10,3
        uh = np.einsum("qa,cak->cqk", basis, coef)
        duh = np.einsum("cqam,cak->cqkm", dbasis, coef)
    else:
        coef = approx.values()[dm.cell_p1]                              # nC,4
        uh = np.einsum("qa,ca->cq", lam, coef)
        duh = np.einsum("cam,ca->cm", grads, coef)[:, None, :].repeat(nq, 1)
    ue = exact.eval(X.reshape(-1, 3), t).reshape(uh.shape)
    err = (ue - uh) ** 2
    err = err.reshape(nC, nq, -1).sum(axis=2)
    total = np.einsum("cq,q,c->", err, qw, det)
    if norm_type.upper() == "H1":
        h = 1e-6
        ge = []
        for m in range(3):
            Xp = X.reshape(-1, 3).copy()
            Xm = Xp.copy()
            Xp[:, m] += h
            Xm[:, m] -= h
            ge.append((exact.eval(Xp, t) - exact.eval(Xm, t)) / (2 * h))
        ge = np.stack(ge, axis=-1).reshape(duh.shape)
        e2 = ((ge - duh) ** 2).reshape(nC, nq, -1).sum(axis=2)
        total += np.einsum("cq,q,c->", e2, qw, det)
    return math.sqrt(total)


def p2_values(lam):
    out = [lam[:, i] * (2 * lam[:, i] - 1) for i in range(4)]
    out += [4 * lam[:, i] * lam[:, j] for (i, j) in refel.EDGES]
    return np.stack(out, axis=1)


def p2_grads(lam, grads):
    """(nC, nq, 10, 3) physical gradients of the P2 basis."""
    out = [(4 * lam[None, :, i, None] - 1) * grads[:, None, i, :] for i in range(4)]
    out += [4 * (lam[None, :, j, None] * grads[:, None, i, :]
                 + lam[None, :, i, None] * grads[:, None, j, :])
            for (i, j) in refel.EDGES]
    return np.stack(out, axis=2)


def tet_quadrature(order):
    """Collapsed Gauss-Jacobi (Stroud conical product) rule on the reference
    tetrahedron, exact for polynomials of total degree <= order."""
    from scipy.special import roots_jacobi
    n = order // 2 + 1
    a0, w0 = roots_jacobi(n, 0, 0)
    a1, w1 = roots_jacobi(n, 1, 0)
    a2, w2 = roots_jacobi(n, 2, 0)
    pts, wts = [], []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                r, s, t = a0[i], a1[j], a2[k]
                z = (1 + t) / 2
                y = (1 + s) / 2 * (1 - z)
                xx = (1 + r) / 2 * (1 - y - z)
                pts.append((xx, y, z))
                wts.append(w0[i] * w1[j] * w2[k] / 64.0)
    return np.array(pts), np.array(wts)
