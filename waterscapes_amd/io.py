"""Mesh / marker input and per-step field output without an HDF5 library.

The reference reads meshes and facet markers from FEniCS HDF5 files
(`HDF5File.read(mesh, "/mesh")`, `read(boundaries, "/boundaries")`,
sandbox/brain-simulations/MPET-4networks-colin27.py:25-47; test_donut.py:
37-40) and writes every step's displacement and network pressures twice:
to HDF5 time series (`HDF5File.write(u, "/u", t)`, :197-199, :257-259) and
to ParaView collections (`File("u.pvd") << u`, :203-205, :262-264).  The
image has no HDF5 library, so this module keeps the same roles with formats
that need none:

  * save_mesh / load_mesh: coordinates, cells and one facet marker array
    per MeshFunction in a NumPy .npz archive (facets stored by their three
    vertices, so the markers survive any renumbering of the facets);
  * XDMFFile: XDMF 3 time series whose data items are raw little-endian
    binary in one side file (read by ParaView / VisIt / meshio), one
    Grid per write(field, t), sharing geometry and topology;
  * File: "name.pvd" collections of VTK XML unstructured grids (appended
    raw binary), one .vtu per `<<`.
A P2 displacement is written on its own nodes as quadratic tetrahedra
(XDMF Tetrahedron_10 / VTK type 24), P1 fields on the vertices.
"""
from __future__ import annotations

import os
import xml.etree.ElementTree as ET

import numpy as np

from .fem import Function, FieldView, MixedSpace
from .mesh import Mesh, MeshFunction
from . import refel

# the P2 nodes of a cell are its vertices, then its edges in UFC order
# (refel.EDGES); VTK / XDMF quadratic tetrahedra list the edges (0,1),
# (1,2), (0,2), (0,3), (1,3), (2,3)
_VTK_EDGES = ((0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3))
_P2_TO_VTK = [0, 1, 2, 3] + [4 + list(refel.EDGES).index(tuple(sorted(e)))
                             for e in _VTK_EDGES]


# ---------------------------------------------------------------------------
# meshes and markers
# ---------------------------------------------------------------------------
def save_mesh(path, mesh, **markers):
    """Write mesh and facet markers (name=MeshFunction of dimension 2) to
    an .npz archive."""
    data = {"coordinates": mesh.coordinates(), "cells": mesh.cells()}
    for name, mf in markers.items():
        if int(getattr(mf, "dim", 2)) != 2:
            raise ValueError("only facet markers (dimension 2) are stored")
        data[f"facets_{name}"] = mesh.facets
        data[f"values_{name}"] = np.asarray(mf.array())
    np.savez(path, **data)


def load_mesh(path):
    """(mesh, {name: MeshFunction}) from an archive written by save_mesh
    (or any .npz with `coordinates` (nV, 3) and `cells` (nC, 4), plus
    `facets_<name>` (m, 3) vertex triplets and `values_<name>` (m,))."""
    with np.load(path, allow_pickle=False) as z:
        mesh = Mesh(z["coordinates"], z["cells"])
        nV = mesh.num_vertices()
        own = np.sort(mesh.facets, axis=1)
        key_own = (own[:, 0] * nV + own[:, 1]) * nV + own[:, 2]
        order = np.argsort(key_own)
        out = {}
        for k in z.files:
            if not k.startswith("facets_"):
                continue
            name = k[len("facets_"):]
            tri = np.sort(np.asarray(z[k], dtype=np.int64), axis=1)
            vals = np.asarray(z[f"values_{name}"])
            key = (tri[:, 0] * nV + tri[:, 1]) * nV + tri[:, 2]
            pos = np.searchsorted(key_own[order], key)
            pos = np.minimum(pos, key_own.size - 1)
            if not np.array_equal(key_own[order][pos], key):
                raise ValueError(f"markers '{name}' name facets that are not in the mesh")
            mf = MeshFunction("size_t", mesh, 2, 0)
            arr = np.zeros(mesh.num_facets(), dtype=vals.dtype)
            arr[order[pos]] = vals
            mf.array()[:] = arr
            out[name] = mf
    return mesh, out


# ---------------------------------------------------------------------------
# fields
# ---------------------------------------------------------------------------
def _field(obj):
    """(values (n, ncomp), nodes coordinates, cells in VTK order, degree,
    name) of up.split()[i], up.sub(i) or a Function on one field."""
    if isinstance(obj, FieldView):
        space = obj.parent.function_space()
        i = obj.i
        vals = obj.values()
    elif isinstance(obj, Function) and not isinstance(obj.function_space(), MixedSpace):
        cs = obj.function_space()
        space, i = cs.parent, cs.i
        vals = obj.field_values(i)
    else:
        raise TypeError("write a field: up.split()[i], up.sub(i) or a collapsed Function")
    if i > space.F:
        raise ValueError("nullspace multipliers are not a field on the mesh")
    dm = space.dofmap
    if i == 0:
        X = dm.x2
        cells = np.asarray(dm.cell_p2)[:, _P2_TO_VTK]
        return np.asarray(vals, float).reshape(-1, 3), X, cells, 2
    X = space.mesh.coordinates()
    cells = np.asarray(dm.cell_p1)
    return np.asarray(vals, float).reshape(-1, 1), X, cells, 1


class XDMFFile:
    """Time series of one field (like HDF5File(..., "w").write(u, "/u", t)):
    write(field, t) appends the values to `<name>.bin` and rewrites
    `<name>.xdmf`.  Geometry and topology are written once."""

    def __init__(self, path, name=None):
        if not path.endswith(".xdmf"):
            raise ValueError("XDMFFile path must end in .xdmf")
        self.path = path
        self.bin = path[:-5] + ".bin"
        self.name = name or os.path.splitext(os.path.basename(path))[0]
        self.steps = []
        self.mesh_items = None
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        open(self.bin, "wb").close()

    def _append(self, arr):
        arr = np.ascontiguousarray(arr)
        with open(self.bin, "ab") as f:
            off = f.tell()
            f.write(arr.astype(arr.dtype.newbyteorder("<"), copy=False).tobytes())
        return off

    def write(self, field, t):
        vals, X, cells, deg = _field(field)
        if self.mesh_items is None:
            self.mesh_items = (self._append(X.astype(np.float64)), X.shape,
                               self._append(cells.astype(np.int64)), cells.shape, deg)
        elif self.mesh_items[1] != X.shape or self.mesh_items[3] != cells.shape:
            raise ValueError("XDMFFile: the field lives on another mesh than earlier steps")
        off = self._append(vals.astype(np.float64))
        self.steps.append((float(t), off, vals.shape))
        self._write_xml()

    def _item(self, parent, off, shape, ntype, prec):
        d = ET.SubElement(parent, "DataItem", Format="Binary", NumberType=ntype,
                          Precision=str(prec), Endian="Little", Seek=str(off),
                          Dimensions=" ".join(str(s) for s in shape))
        d.text = os.path.basename(self.bin)

    def _write_xml(self):
        xoff, xshape, coff, cshape, deg = self.mesh_items
        root = ET.Element("Xdmf", Version="3.0")
        dom = ET.SubElement(root, "Domain")
        coll = ET.SubElement(dom, "Grid", Name=self.name, GridType="Collection",
                             CollectionType="Temporal")
        for t, off, shape in self.steps:
            g = ET.SubElement(coll, "Grid", Name=f"{self.name}_{t:.12g}", GridType="Uniform")
            ET.SubElement(g, "Time", Value=f"{t:.17g}")
            topo = ET.SubElement(g, "Topology", NumberOfElements=str(cshape[0]),
                                 TopologyType="Tetrahedron_10" if deg == 2 else "Tetrahedron")
            self._item(topo, coff, cshape, "Int", 8)
            geo = ET.SubElement(g, "Geometry", GeometryType="XYZ")
            self._item(geo, xoff, xshape, "Float", 8)
            att = ET.SubElement(g, "Attribute", Name=self.name, Center="Node",
                                AttributeType="Vector" if shape[1] == 3 else "Scalar")
            self._item(att, off, shape if shape[1] == 3 else (shape[0],), "Float", 8)
        ET.indent(root)
        tmp = self.path + ".tmp"
        ET.ElementTree(root).write(tmp, xml_declaration=True, encoding="utf-8")
        os.replace(tmp, self.path)

    def close(self):
        pass


def _vtu(path, vals, X, cells, deg, name):
    """VTK XML UnstructuredGrid with appended raw binary arrays."""
    nC = cells.shape[0]
    k = cells.shape[1]
    arrays = [("points", X.astype("<f8")), ("connectivity", cells.reshape(-1).astype("<i8")),
              ("offsets", (k * np.arange(1, nC + 1)).astype("<i8")),
              ("types", np.full(nC, 24 if deg == 2 else 10, dtype=np.uint8)),
              (name, vals.astype("<f8"))]
    offs, pos = {}, 0
    for nm, a in arrays:
        offs[nm] = pos
        pos += 8 + a.nbytes
    ncomp = vals.shape[1]
    head = (
        '<?xml version="1.0"?>\n'
        '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" '
        'header_type="UInt64">\n<UnstructuredGrid>\n'
        f'<Piece NumberOfPoints="{X.shape[0]}" NumberOfCells="{nC}">\n'
        f'<PointData><DataArray type="Float64" Name="{name}" NumberOfComponents="{ncomp}" '
        f'format="appended" offset="{offs[name]}"/></PointData>\n'
        '<Points><DataArray type="Float64" NumberOfComponents="3" format="appended" '
        f'offset="{offs["points"]}"/></Points>\n<Cells>\n'
        f'<DataArray type="Int64" Name="connectivity" format="appended" '
        f'offset="{offs["connectivity"]}"/>\n'
        f'<DataArray type="Int64" Name="offsets" format="appended" offset="{offs["offsets"]}"/>\n'
        f'<DataArray type="UInt8" Name="types" format="appended" offset="{offs["types"]}"/>\n'
        '</Cells>\n</Piece>\n</UnstructuredGrid>\n<AppendedData encoding="raw">\n_')
    with open(path, "wb") as f:
        f.write(head.encode())
        for nm, a in arrays:
            f.write(np.uint64(a.nbytes).tobytes())
            f.write(np.ascontiguousarray(a).tobytes())
        f.write(b"\n</AppendedData>\n</VTKFile>\n")


class File:
    """ParaView collection `name.pvd` (dolfin File): `file << field` or
    `file << (field, t)` writes `name{k:06d}.vtu` and rewrites the .pvd."""

    def __init__(self, path):
        if not path.endswith(".pvd"):
            raise ValueError("File path must end in .pvd")
        self.path = path
        self.stem = os.path.splitext(os.path.basename(path))[0]
        self.dir = os.path.dirname(os.path.abspath(path))
        os.makedirs(self.dir, exist_ok=True)
        self.entries = []

    def __lshift__(self, item):
        field, t = item if isinstance(item, tuple) else (item, float(len(self.entries)))
        vals, X, cells, deg = _field(field)
        fn = f"{self.stem}{len(self.entries):06d}.vtu"
        _vtu(os.path.join(self.dir, fn), vals, X, cells, deg, self.stem)
        self.entries.append((float(t), fn))
        root = ET.Element("VTKFile", type="Collection", version="0.1")
        col = ET.SubElement(root, "Collection")
        for tt, f in self.entries:
            ET.SubElement(col, "DataSet", timestep=f"{tt:.17g}", part="0", file=f)
        ET.indent(root)
        ET.ElementTree(root).write(self.path, xml_declaration=True, encoding="utf-8")
        return self
