"""Mesh / marker archives and per-step field output (waterscapes_amd/io.py),
the library-free counterparts of the reference's HDF5File mesh input and
HDF5 / PVD output (sandbox/brain-simulations/MPET-4networks-colin27.py:
25-47, :197-205, :257-264).  CPU only: fields are host Functions."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from waterscapes_amd import io
from waterscapes_amd.fem import Function, MixedSpace
from waterscapes_amd.mesh import CompiledSubDomain, MeshFunction, UnitCubeMesh


def _state(n=2, A=2, seed=0):
    mesh = UnitCubeMesh(n, n, n)
    space = MixedSpace(mesh, A)
    rng = np.random.default_rng(seed)
    return mesh, space, Function(space, rng.standard_normal(space.dim()))


def test_mesh_archive_round_trip_survives_renumbering(tmp_path):
    mesh = UnitCubeMesh(3, 2, 2)
    mk = MeshFunction("size_t", mesh, 2, 7)
    CompiledSubDomain("on_boundary && near(x[2], 0.0)").mark(mk, 0)
    CompiledSubDomain("on_boundary && near(x[0], 1.0)").mark(mk, 2)
    path = str(tmp_path / "m.npz")
    io.save_mesh(path, mesh, boundaries=mk)
    m2, marks = io.load_mesh(path)
    assert np.array_equal(m2.coordinates(), mesh.coordinates())
    assert np.array_equal(marks["boundaries"].array(), mk.array())
    # the same mesh with permuted vertices and cells: markers follow the facets
    rng = np.random.default_rng(1)
    pv = rng.permutation(mesh.num_vertices())
    inv = np.empty_like(pv)
    inv[pv] = np.arange(pv.size)
    coords = mesh.coordinates()[pv]
    cells = inv[mesh.cells()][rng.permutation(mesh.num_cells())]
    z = dict(np.load(path))
    z["coordinates"], z["cells"] = coords, cells
    z["facets_boundaries"] = inv[z["facets_boundaries"]]
    p2 = str(tmp_path / "perm.npz")
    np.savez(p2, **z)
    m3, marks3 = io.load_mesh(p2)
    mid_a = mesh.facet_midpoints()
    mid_b = m3.facet_midpoints()
    ka = np.lexsort(np.round(mid_a, 12).T)
    kb = np.lexsort(np.round(mid_b, 12).T)
    assert np.array_equal(mk.array()[ka], marks3["boundaries"].array()[kb])


def _read_item(xdmf_path, item):
    base = os.path.dirname(xdmf_path)
    shape = tuple(int(s) for s in item.get("Dimensions").split())
    dt = {"Float": "<f8", "Int": "<i8"}[item.get("NumberType")]
    with open(os.path.join(base, item.text.strip()), "rb") as f:
        f.seek(int(item.get("Seek")))
        return np.frombuffer(f.read(int(np.prod(shape)) * 8), dtype=dt).reshape(shape)


def test_xdmf_time_series_round_trip(tmp_path):
    mesh, space, up = _state()
    dm = space.dofmap
    fu = io.XDMFFile(str(tmp_path / "u.xdmf"))
    fp = io.XDMFFile(str(tmp_path / "p1.xdmf"))
    states = []
    for k, t in enumerate((0.1, 0.2, 0.3)):
        up.data[:] = np.random.default_rng(k).standard_normal(space.dim())
        states.append(up.data.copy())
        parts = up.split()
        fu.write(parts[0], t)
        fp.write(parts[2], t)
    root = ET.parse(str(tmp_path / "u.xdmf")).getroot()
    grids = root.findall("./Domain/Grid/Grid")
    assert [float(g.find("Time").get("Value")) for g in grids] == [0.1, 0.2, 0.3]
    for g, st in zip(grids, states):
        topo = g.find("Topology")
        assert topo.get("TopologyType") == "Tetrahedron_10"
        cells = _read_item(str(tmp_path / "u.xdmf"), topo.find("DataItem"))
        X = _read_item(str(tmp_path / "u.xdmf"), g.find("Geometry/DataItem"))
        vals = _read_item(str(tmp_path / "u.xdmf"), g.find("Attribute/DataItem"))
        assert np.array_equal(X, dm.x2)
        assert np.array_equal(vals, st[: space.nu].reshape(-1, 3))
        # quadratic-tetrahedron node order: edge node k is the midpoint of
        # the k-th VTK edge of the cell's first four nodes
        for k, (a, b) in enumerate(((0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3))):
            assert np.allclose(X[cells[:, 4 + k]], 0.5 * (X[cells[:, a]] + X[cells[:, b]]))
    rootp = ET.parse(str(tmp_path / "p1.xdmf")).getroot()
    g = rootp.findall("./Domain/Grid/Grid")[1]
    vals = _read_item(str(tmp_path / "p1.xdmf"), g.find("Attribute/DataItem"))
    assert np.array_equal(vals, states[1][space.field_index(2)])
    assert g.find("Topology").get("TopologyType") == "Tetrahedron"


def test_pvd_collection(tmp_path):
    mesh, space, up = _state()
    f = io.File(str(tmp_path / "pvd" / "p0.pvd"))
    f << (up.split()[1], 0.5)
    f << (up.split()[1], 1.0)
    root = ET.parse(str(tmp_path / "pvd" / "p0.pvd")).getroot()
    ds = root.findall("./Collection/DataSet")
    assert [d.get("timestep") for d in ds] == ["0.5", "1"]
    vtu = tmp_path / "pvd" / ds[0].get("file")
    raw = vtu.read_bytes()
    head, data = raw.split(b'<AppendedData encoding="raw">\n_', 1)
    hx = ET.fromstring(head.decode() + "<AppendedData encoding=\"raw\"/></VTKFile>")
    piece = hx.find("./UnstructuredGrid/Piece")
    assert int(piece.get("NumberOfPoints")) == mesh.num_vertices()
    assert int(piece.get("NumberOfCells")) == mesh.num_cells()
    arr = piece.find("./PointData/DataArray")
    off = int(arr.get("offset"))
    nb = int(np.frombuffer(data[off:off + 8], "<u8")[0])
    vals = np.frombuffer(data[off + 8: off + 8 + nb], "<f8")
    assert np.array_equal(vals, up.data[space.field_index(1)])


def test_field_output_rejects_multipliers_and_mixed_functions(tmp_path):
    mesh, space, up = _state()
    with pytest.raises(TypeError):
        io.XDMFFile(str(tmp_path / "x.xdmf")).write(up, 0.0)
    sp2 = MixedSpace(mesh, 2, u_nullspace=True)
    f2 = Function(sp2, np.zeros(sp2.dim()))
    with pytest.raises(ValueError):
        io.XDMFFile(str(tmp_path / "y.xdmf")).write(f2.split()[sp2.F + 1], 0.0)
