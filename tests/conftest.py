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
