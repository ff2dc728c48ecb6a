"""Packing of instance hierarchies into one block-diagonal batch hierarchy
(CPU, host arrays).  A V-cycle on the packed hierarchy must act on each
instance's block like that instance's own V-cycle on (s_k A_k): padding
levels (identity prolongators ending in the instance's own coarsest solve)
and the rho / scale folding into the diagonal may change rounding only."""
import numpy as np
import pytest
import scipy.sparse as sp

from test_gpu_solver import _ref_vcycle


def _laplacian(n, shift=0.0):
    e = np.ones(n)
    T = sp.diags([-e[:-1], 2 * e, -e[:-1]], [-1, 0, 1])
    I = sp.identity(n)
    L = (sp.kron(sp.kron(T, I), I) + sp.kron(sp.kron(I, T), I) + sp.kron(sp.kron(I, I), T)
         + shift * sp.identity(n ** 3)).tocsr()
    L.sort_indices()
    return L


@pytest.mark.parametrize("deg,levdeg", [(2, ()), (3, (1,)), (2, (1, 3))])
def test_packed_vcycle_equals_blockwise(deg, levdeg):
    from waterscapes_amd import amg
    from waterscapes_amd.amg import AmgParams
    from waterscapes_amd.batch import pack_hierarchy
    prm = AmgParams(max_coarse=40, smoother_degree=deg, level_degrees=levdeg)
    mats = [_laplacian(4, 0.5), _laplacian(9), _laplacian(6, 0.1)]
    hiers = [amg.setup_host(A.indptr, A.indices, A.data, prm) for A in mats]
    depths = [len(h[0]) for h in hiers]
    assert len(set(depths)) > 1, depths          # padding is exercised
    scales = [2.0, 0.5, 7.0]
    levels, (inv, blk, off, ioff) = pack_hierarchy(hiers, scales)
    assert len(levels) == max(depths)
    # dense coarse blocks -> one block-diagonal inverse for the restatement
    sizes = np.diff(off)
    Cinv = sp.block_diag([inv[ioff[k]:ioff[k + 1]].reshape(n, n) for k, n in enumerate(sizes)]
                         ).toarray()
    assert np.array_equal(blk, np.repeat(np.arange(3), sizes))
    A0 = sp.csr_matrix((levels[0]["A_v"], levels[0]["A_ci"], levels[0]["A_rp"]),
                       shape=(levels[0]["n"],) * 2)
    rng = np.random.default_rng(1)
    b = rng.standard_normal((A0.shape[0], 1))
    z = _ref_vcycle(levels, Cinv, A0, None, prm, b)
    r0 = 0
    for k, (A, (lv, ci)) in enumerate(zip(mats, hiers)):
        n = A.shape[0]
        zk = _ref_vcycle(lv, ci, A, None, prm, b[r0:r0 + n]) / scales[k]
        err = np.abs(z[r0:r0 + n] - zk).max() / np.abs(zk).max()
        assert err < 1e-12, (k, err)
        r0 += n
    assert r0 == A0.shape[0]
    # grp: the instance of every row of every level, and no nonzero of A, P
    # or R couples two instances (what lets the device V-cycle skip the rows
    # of instances that stopped)
    for l, lev in enumerate(levels):
        rows = [lv[0][min(l, len(lv[0]) - 1)]["n"] for lv in hiers]
        assert np.array_equal(lev["grp"], np.repeat(np.arange(3), rows)), l
        g = lev["grp"]
        A = sp.csr_matrix((lev["A_v"], lev["A_ci"], lev["A_rp"]), shape=(lev["n"],) * 2).tocoo()
        assert np.array_equal(g[A.row], g[A.col]), l
        if lev["nc"]:
            gc = levels[l + 1]["grp"]
            P = sp.csr_matrix((lev["P_v"], lev["P_ci"], lev["P_rp"]),
                              shape=(lev["n"], lev["nc"])).tocoo()
            assert np.array_equal(g[P.row], gc[P.col]), l
            R = sp.csr_matrix((lev["R_v"], lev["R_ci"], lev["R_rp"]),
                              shape=(lev["nc"], lev["n"])).tocoo()
            assert np.array_equal(gc[R.row], g[R.col]), l
    # so changing one instance's right-hand side (here to huge values) leaves
    # the others' results bit-identical
    bn = b.copy()
    sel = levels[0]["grp"] == 1
    bn[sel] = 1e30 * rng.standard_normal((int(sel.sum()), 1))
    zn = _ref_vcycle(levels, Cinv, A0, None, prm, bn)
    other = levels[0]["grp"] != 1
    assert np.array_equal(zn[other], z[other])


def test_csr_blockdiag_matches_scipy():
    from waterscapes_amd.batch import csr_blockdiag
    rng = np.random.default_rng(2)
    mats = [sp.random(5, 3, 0.4, random_state=1, format="csr"),
            sp.random(2, 6, 0.5, random_state=2, format="csr"),
            sp.random(4, 4, 0.3, random_state=3, format="csr")]
    rp, ci, v, nr, nc = csr_blockdiag([(M.indptr, M.indices, M.data, M.shape[0], M.shape[1])
                                       for M in mats])
    got = sp.csr_matrix((v, ci, rp), shape=(nr, nc)).toarray()
    ref = sp.block_diag(mats).toarray()
    assert np.array_equal(got, ref)
    del rng
