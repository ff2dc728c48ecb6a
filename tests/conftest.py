import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmpet_hip.so)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:  # pragma: no cover
        return False


def pytest_collection_modifyitems(config, items):
    if has_gpu():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


def oracle_mesh(space):
    """Raw arrays the oracle consumes, taken from a product MixedSpace (the
    numbering is an input of the oracle; everything else it recomputes)."""
    mesh = space.mesh
    dm = space.dofmap
    return dict(coords=mesh.coordinates(), cells=dm.cell_p1.astype(np.int64),
                cell_p2=dm.cell_p2.astype(np.int64), cell_p1=dm.cell_p1.astype(np.int64),
                nP2=dm.nP2, nP1=dm.nP1, x2=dm.x2, facets=mesh.facets,
                facet_cell=mesh.facet_cell, facet_local=mesh.facet_local,
                exterior=mesh.exterior, cell_facets=mesh.cell_facets)


CFG1 = dict(J=2, E=1e4, nu=0.35, alpha=(0.5, 0.5), c=(1e-4, 1e-4), K=(1e-3, 1e-4),
            S=((0.0, 1e-2), (1e-2, 0.0)))


def cfg2_params(A=4):
    med = (1e-6, 1e-7, 1e-6, 1e-5, 1e-6, 1e-6, 1e-6)[:A]
    S = [[1e-3 if abs(i - j) == 1 else 0.0 for j in range(A)] for i in range(A)]
    return dict(J=A, E=1500.0, nu=0.479, alpha=(0.25,) * A, c=(1e-5,) * A, K=med, S=S)


def lognormal_K(nC, medians, sigma=0.5, seed=0):
    rng = np.random.default_rng(seed)
    med = np.asarray(medians, float)
    return med[None, :] * np.exp(sigma * rng.standard_normal((nC, med.size)))


def sampled_row_residuals(m, params, K_cells, dt, theta, x, x_prev, rows):
    """Residual of the oracle's equations (A x)_i = (B x_prev)_i at the given
    rows (rows with no Dirichlet condition; a workload with no sources),
    with A and B assembled by the oracle from only the cells that touch
    those rows: returns (r, scale), scale_i = sum_j |A_ij x_j| + sum_j
    |B_ij x_prev_j|.  Lets a solution of any size be checked against the
    oracle's own assembly at a sample of its rows."""
    from oracle import mpet_oracle as O
    rows = np.asarray(rows, dtype=np.int64)
    lay = O.Layout(m["nP2"], m["nP1"], int(params["J"]))
    n2 = np.unique(rows[rows < lay.nu] // 3)
    n1 = np.unique((rows[rows >= lay.nu] - lay.nu) // lay.F)
    sel = np.isin(m["cell_p2"], n2).any(1) | np.isin(m["cell_p1"], n1).any(1)
    c = np.nonzero(sel)[0]
    Kc = None if K_cells is None else np.asarray(K_cells)[c]
    args = (m["coords"], m["cells"][c], m["cell_p2"][c], m["cell_p1"][c], m["nP2"], m["nP1"],
            params, dt, theta, Kc)
    A = O.assemble_matrix(*args, which="A")[rows]
    B = O.rhs_matrix(m["coords"], m["cells"][c], m["cell_p1"][c], m["nP2"], m["nP1"], params,
                     dt, theta, Kc)[rows]
    r = A @ x - B @ x_prev
    scale = abs(A) @ np.abs(x) + abs(B) @ np.abs(x_prev)
    return r, scale


def field_row_ratios(lay, rows, r, scale):
    """||r|| / ||scale|| over the sampled rows of each field (u, p_0 .. p_A)."""
    rows = np.asarray(rows)
    fld = np.where(rows < lay.nu, 0, 1 + (rows - lay.nu) % lay.F)
    return [float(np.linalg.norm(r[fld == k]) / np.linalg.norm(scale[fld == k]))
            for k in range(lay.F + 1) if np.any(fld == k)]
