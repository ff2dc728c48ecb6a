"""Host smoothed-aggregation setup (csrc/amg_setup.cpp) on the CPU: the
dense inverse of the coarsest operator is the inverse (A_c Cinv = I,
symmetric) and does not depend on the OpenMP thread count; the levels are
Galerkin products R A P with R = P^T."""
import os
import subprocess
import sys

import numpy as np
import scipy.sparse as sp

from waterscapes_amd import amg


def _lap3d(n):
    """7-point Laplacian on an n^3 grid plus a small shift (SPD)."""
    e = np.ones(n)
    T = sp.diags([-e[:-1], 2 * e, -e[:-1]], [-1, 0, 1])
    I = sp.identity(n)
    A = sp.kron(sp.kron(T, I), I) + sp.kron(sp.kron(I, T), I) + sp.kron(sp.kron(I, I), T)
    return (A + 1e-3 * sp.identity(n ** 3)).tocsr()


def test_coarse_inverse_and_galerkin_levels():
    A = _lap3d(14)
    A.sort_indices()
    levels, cinv = amg.setup_host(A.indptr, A.indices, A.data,
                                  amg.AmgParams(max_coarse=700, max_levels=8))
    assert len(levels) >= 2
    last = levels[-1]
    Ac = sp.csr_matrix((last["A_v"], last["A_ci"], last["A_rp"]), shape=(last["n"],) * 2)
    I = Ac.toarray() @ cinv
    assert np.abs(I - np.eye(last["n"])).max() <= 1e-10
    assert np.array_equal(cinv, cinv.T)
    for l in range(len(levels) - 1):
        lev, nxt = levels[l], levels[l + 1]
        Al = sp.csr_matrix((lev["A_v"], lev["A_ci"], lev["A_rp"]), shape=(lev["n"],) * 2)
        P = sp.csr_matrix((lev["P_v"], lev["P_ci"], lev["P_rp"]), shape=(lev["n"], lev["nc"]))
        An = sp.csr_matrix((nxt["A_v"], nxt["A_ci"], nxt["A_rp"]), shape=(nxt["n"],) * 2)
        G = (P.T @ Al @ P).toarray()
        assert np.abs(G - An.toarray()).max() <= 1e-12 * np.abs(G).max()


def test_coarse_inverse_independent_of_thread_count():
    code = ("import numpy as np, scipy.sparse as sp, sys; sys.path.insert(0, %r);"
            "from tests.test_amg_setup import _lap3d; from waterscapes_amd import amg;"
            "A = _lap3d(12); A.sort_indices();"
            "lv, ci = amg.setup_host(A.indptr, A.indices, A.data, amg.AmgParams(max_coarse=900));"
            "sys.stdout.buffer.write(ci.tobytes())") % os.path.dirname(os.path.dirname(
                os.path.abspath(__file__)))
    outs = []
    for t in ("1", "4"):
        env = dict(os.environ, OMP_NUM_THREADS=t)
        outs.append(subprocess.run([sys.executable, "-c", code], env=env, check=True,
                                   capture_output=True).stdout)
    assert len(outs[0]) > 0 and outs[0] == outs[1]
