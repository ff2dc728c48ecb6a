"""CPU check of the cell-wise operator's local formulas (csrc/cellwise.h).

The kernel integrates 2 mu eps(u):eps(v) + p0 div v and div(u) w0 on each
tetrahedron from the vertex values of grad u (exact: P2 gradients are
linear).  Restated here in FP64 numpy on single cells, the local results
must equal the sympy-exact element tensors of the golden fixtures
(E = int eps:eps, B = int psi div phi): y_u = 2 mu E u + B p0,
y_p0 = B^T u, for random local vectors.  A wrong edge pairing, a dropped
term of the vertex gradient, the factor 1/20 vs 1/5, or a transposed
stress would all fail it."""
import os

import numpy as np
import pytest

from waterscapes_amd import refel

GOLD = os.path.join(os.path.dirname(__file__), "golden", "element_tensors.npz")


def _edge(j, k):
    return refel.EDGES.index((min(j, k), max(j, k)))


def cellwise_local(X, u, p0, mu):
    """Local u results (10, 3) and p0 results (4,) of one tetrahedron with
    vertices X (4, 3), local P2 values u (10, 3) and P1 p0 (4,)."""
    J = np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]], axis=1)
    Gi = np.linalg.inv(J)
    G = np.vstack([-Gi.sum(0), Gi])            # rows: grad lambda_j
    vol = abs(np.linalg.det(J)) / 6.0
    U = sum(np.outer(u[j], G[j]) for j in range(4))
    s, tr = [], []
    for n in range(4):
        g = 4.0 * np.outer(u[n], G[n]) - U
        for k in range(4):
            if k != n:
                g += 4.0 * np.outer(u[4 + _edge(n, k)], G[k])
        tr.append(np.trace(g))
        s.append(mu * (g + g.T) + p0[n] * np.eye(3))
    S = sum(s)
    c = vol / 20.0
    y = np.zeros((10, 3))
    for j in range(4):
        y[j] = c * (4.0 * s[j] - S) @ G[j]
    for m, (j, k) in enumerate(refel.EDGES):
        y[4 + m] = 4.0 * c * ((S + s[k]) @ G[j] + (S + s[j]) @ G[k])
    D = sum(tr)
    yp = np.array([c * (D + tr[v]) for v in range(4)])
    return y, yp


@pytest.mark.parametrize("k", range(4))
def test_cellwise_local_equals_exact_element_tensors(k):
    d = np.load(GOLD)
    X = d["coords"][k]
    E, B = d[f"E{k}"], d[f"B{k}"]
    rng = np.random.default_rng(10 + k)
    mu = 1.7
    for _ in range(3):
        u = rng.standard_normal((10, 3))
        p0 = rng.standard_normal(4)
        y, yp = cellwise_local(X, u, p0, mu)
        ref_u = 2.0 * mu * E @ u.reshape(-1) + B @ p0
        ref_p = B.T @ u.reshape(-1)
        scale = np.abs(ref_u).max()
        assert np.abs(y.reshape(-1) - ref_u).max() <= 1e-12 * scale
        assert np.abs(yp - ref_p).max() <= 1e-12 * np.abs(ref_p).max()


def test_cellwise_edge_indexing_matches_kernel():
    """The kernel's closed-form edge index (cw_edge) against the UFC table."""
    def cw_edge(j, k):
        lo, hi = min(j, k), max(j, k)
        return 6 - hi if lo == 0 else (4 - hi if lo == 1 else 0)
    for m, (j, k) in enumerate(refel.EDGES):
        assert cw_edge(j, k) == m and cw_edge(k, j) == m
