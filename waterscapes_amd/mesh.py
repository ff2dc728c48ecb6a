"""Tetrahedral meshes, facet markers and subdomains.

Minimal look-alikes of the DOLFIN objects the reference hot path consumes:
``Mesh`` (vertices + cells), ``UnitCubeMesh`` / ``BoxMesh`` (6 tetrahedra per
cube), ``MeshFunction("size_t", mesh, 2)`` facet markers with ``set_all`` /
``array``, and ``SubDomain`` / ``CompiledSubDomain`` with ``mark``.  They follow
the semantics the reference relies on:
  * markers default to INVALID = sys.maxsize (mpet/mpetproblem.py:153-163);
  * a facet is marked when the predicate holds at its midpoint and at all its
    vertices, ``on_boundary`` being true for exterior facets (DOLFIN
    SubDomain.mark semantics used by the reference tests,
    test_convergence_totalpressuresolver.py:145-154).
"""
from __future__ import annotations

import re
import sys

import numpy as np

from . import refel

DOLFIN_EPS = 3.0e-16
INVALID = sys.maxsize


class Mesh:
    """Affine tetrahedral mesh.  ``coordinates`` (nV, 3) float64 and ``cells``
    (nC, 4) int64 with each row sorted ascending (UFC local ordering)."""

    def __init__(self, coordinates, cells, structured=None):
        coords = np.ascontiguousarray(coordinates, dtype=np.float64)
        cells = np.sort(np.asarray(cells, dtype=np.int64), axis=1)
        if coords.ndim != 2 or coords.shape[1] != 3:
            raise NotImplementedError("only 3D tetrahedral meshes are supported")
        if cells.ndim != 2 or cells.shape[1] != 4:
            raise ValueError("cells must be (nC, 4) vertex indices")
        self._coords = coords
        self._cells = np.ascontiguousarray(cells)
        self.structured = structured          # (nx, ny, nz) for box meshes
        self._build_topology()

    # -- DOLFIN-like accessors ------------------------------------------------
    def coordinates(self):
        return self._coords

    def cells(self):
        return self._cells

    def num_vertices(self):
        return self._coords.shape[0]

    def num_cells(self):
        return self._cells.shape[0]

    def num_facets(self):
        return self.facets.shape[0]

    def num_edges(self):
        return self.edges.shape[0]

    class _Geometry:
        def __init__(self, d):
            self._d = d

        def dim(self):
            return self._d

    class _Topology:
        def __init__(self, d):
            self._d = d

        def dim(self):
            return self._d

    def geometry(self):
        return Mesh._Geometry(3)

    def topology(self):
        return Mesh._Topology(3)

    def hmin(self):
        c = self._cells
        x = self._coords
        best = np.inf
        for (i, j) in refel.EDGES:
            best = min(best, np.linalg.norm(x[c[:, i]] - x[c[:, j]], axis=1).min())
        return float(best)

    def hmax(self):
        c = self._cells
        x = self._coords
        best = 0.0
        for (i, j) in refel.EDGES:
            best = max(best, np.linalg.norm(x[c[:, i]] - x[c[:, j]], axis=1).max())
        return float(best)

    # -- topology ---------------------------------------------------------------
    def _build_topology(self):
        c = self._cells
        nC = c.shape[0]
        nV = self._coords.shape[0]
        # edges: 6 per cell in UFC order, global ids by unique vertex pair
        pairs = np.stack([c[:, [i, j]] for (i, j) in refel.EDGES], axis=1)  # nC,6,2
        key = pairs[..., 0] * nV + pairs[..., 1]
        ukey, inv = np.unique(key.reshape(-1), return_inverse=True)
        self.edges = np.stack([ukey // nV, ukey % nV], axis=1)
        self.cell_edges = inv.reshape(nC, 6)
        # facets: face i opposite vertex i
        tri = np.stack([c[:, list(f)] for f in refel.FACES], axis=1)        # nC,4,3
        fkey = (tri[..., 0] * nV + tri[..., 1]) * nV + tri[..., 2]
        ufk, finv, fcnt = np.unique(fkey.reshape(-1), return_inverse=True,
                                    return_counts=True)
        self.facets = np.stack([ufk // (nV * nV), (ufk // nV) % nV, ufk % nV], axis=1)
        self.cell_facets = finv.reshape(nC, 4)
        self.exterior = fcnt == 1
        # owning (cell, local face) of every facet (first occurrence)
        flat = np.arange(nC * 4)
        order = np.argsort(finv, kind="stable")
        first = np.ones(len(ufk), dtype=bool)
        sorted_f = finv[order]
        starts = np.searchsorted(sorted_f, np.arange(len(ufk)))
        owner = flat[order[starts]]
        self.facet_cell = owner // 4
        self.facet_local = owner % 4
        del first

    def cell_volumes(self):
        x = self._coords
        c = self._cells
        J = np.stack([x[c[:, 1]] - x[c[:, 0]], x[c[:, 2]] - x[c[:, 0]],
                      x[c[:, 3]] - x[c[:, 0]]], axis=2)
        return np.abs(np.linalg.det(J)) / 6.0

    def facet_midpoints(self):
        return self._coords[self.facets].mean(axis=1)


def BoxMesh(p0, p1, nx, ny, nz):
    """Box [p0, p1] split into nx*ny*nz cubes of 6 tetrahedra around the
    v0-v7 diagonal (DOLFIN's BoxMesh split)."""
    p0 = np.asarray(p0, dtype=np.float64)
    p1 = np.asarray(p1, dtype=np.float64)
    xs = np.linspace(p0[0], p1[0], nx + 1)
    ys = np.linspace(p0[1], p1[1], ny + 1)
    zs = np.linspace(p0[2], p1[2], nz + 1)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    coords = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
    sx, sy = 1, nx + 1
    sz = (nx + 1) * (ny + 1)
    v0 = ix * sx + iy * sy + iz * sz
    v1 = v0 + sx
    v2 = v0 + sy
    v3 = v1 + sy
    v4 = v0 + sz
    v5 = v1 + sz
    v6 = v2 + sz
    v7 = v3 + sz
    cells = np.stack([
        np.stack([v0, v1, v3, v7], 1), np.stack([v0, v1, v7, v5], 1),
        np.stack([v0, v5, v7, v4], 1), np.stack([v0, v3, v2, v7], 1),
        np.stack([v0, v6, v4, v7], 1), np.stack([v0, v2, v6, v7], 1)], axis=1)
    # order cells cube-by-cube in lexicographic (z, y, x) cube order
    cube = (iz * ny + iy) * nx + ix
    order = np.argsort(cube, kind="stable")
    cells = cells[order].reshape(-1, 4)
    return Mesh(coords, cells, structured=(nx, ny, nz))


def UnitCubeMesh(nx, ny, nz):
    return BoxMesh((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), nx, ny, nz)


# ---------------------------------------------------------------------------
# markers
# ---------------------------------------------------------------------------


class MeshFunction:
    """``MeshFunction("size_t", mesh, dim)``; only facet functions (dim 2)
    and cell functions (dim 3) are meaningful here."""

    def __init__(self, value_type, mesh, dim, value=None):
        self.mesh = mesh
        self.dim = dim
        n = mesh.num_facets() if dim == 2 else mesh.num_cells()
        dtype = np.uint64 if value_type in ("size_t", "uint") else np.float64
        self._values = np.zeros(n, dtype=dtype)
        if value is not None:
            self.set_all(value)

    def set_all(self, value):
        self._values[:] = value

    def array(self):
        return self._values

    def __getitem__(self, i):
        return self._values[i]

    def __setitem__(self, i, v):
        self._values[i] = v

    def size(self):
        return self._values.shape[0]


def FacetFunction(value_type, mesh, value=None):
    return MeshFunction(value_type, mesh, 2, value)


class SubDomain:
    """Subclass and override ``inside(x, on_boundary)``; ``x`` is an (n, 3)
    array of points and ``on_boundary`` an (n,) bool array, the return value a
    bool array.  Marking follows DOLFIN: a facet is marked when the predicate
    holds at its midpoint and at all three vertices."""

    def inside(self, x, on_boundary):  # pragma: no cover - interface
        raise NotImplementedError

    def mark(self, meshfunction, value):
        mesh = meshfunction.mesh
        if meshfunction.dim == 2:
            onb = mesh.exterior
            ok = self.inside(mesh.facet_midpoints(), onb)
            for k in range(3):
                ok = ok & self.inside(mesh.coordinates()[mesh.facets[:, k]], onb)
        else:
            x = mesh.coordinates()[mesh.cells()]
            onb = np.zeros(mesh.num_cells(), dtype=bool)
            ok = self.inside(x.mean(axis=1), onb)
            for k in range(4):
                ok = ok & self.inside(x[:, k], onb)
        meshfunction.array()[np.asarray(ok, dtype=bool)] = value


def _c_to_numpy(expr):
    """Translate a DOLFIN C++ subdomain/expression string into numpy syntax."""
    s = expr.replace("&&", " & ").replace("||", " | ")
    s = re.sub(r"!(?!=)", " ~", s)
    s = re.sub(r"x\[(\d)\]", r"x[:, \1]", s)
    return s


class CompiledSubDomain(SubDomain):
    """``CompiledSubDomain("on_boundary && near(x[0], 1.0)", **params)``.
    The C++ condition is translated to numpy and evaluated on point arrays."""

    def __init__(self, cpp_code, **params):
        self.code = cpp_code
        self.params = dict(params)
        self._src = _c_to_numpy(cpp_code)

    def inside(self, x, on_boundary):
        def near(a, b, eps=DOLFIN_EPS):
            return np.abs(a - b) < eps

        ns = {"x": x, "on_boundary": np.asarray(on_boundary, dtype=bool),
              "near": near, "DOLFIN_EPS": DOLFIN_EPS, "abs": np.abs,
              "fabs": np.abs, "sqrt": np.sqrt, "pi": np.pi, "M_PI": np.pi,
              "true": True, "false": False}
        ns.update(self.params)
        out = eval(self._src, {"__builtins__": {}}, ns)   # noqa: S307 - user code
        return np.broadcast_to(np.asarray(out, dtype=bool), (x.shape[0],))
