"""CPU checks of the block-partial tables of the cell-wise operator
(waterscapes_amd.device.block_partial_tables), emulating what the kernels
do with them: the cell kernel sums, per partial q of its block, the local
slots lsl[lrp[q] .. lrp[q+1]) and stores the sum at q (block-major); the
row kernel sums the partials pe[e] of a node's entries e in rp[a] ..
rp[a+1].  The result must equal a
direct scatter-add of every (cell, local node) value to its node, every
index must stay in range, and each node's entries must come from distinct
blocks in increasing order."""
import numpy as np
import pytest
import torch

from waterscapes_amd.device import block_partial_tables


def _emulate(cell_nodes, nnode, B, vals):
    t = block_partial_tables(torch.as_tensor(cell_nodes), nnode, B)
    nC, k = cell_nodes.shape
    nb = (nC + B - 1) // B
    bq, lrp, gp, pe = t["bq"].numpy(), t["lrp"].numpy(), t["gp"].numpy(), t["pe"].numpy()
    lsl, rp = t["lsl"].numpy().astype(np.int64), t["rp"].numpy()
    npart = t["npart"]
    assert bq[0] == 0 and bq[-1] == npart and np.all(np.diff(bq) >= 0) and bq.size == nb + 1
    assert lrp[0] == 0 and lrp[-1] == nC * k and np.all(np.diff(lrp) >= 1)
    assert np.array_equal(np.sort(gp), np.arange(npart))
    assert np.array_equal(pe[gp], np.arange(npart))          # inverse permutations
    assert rp[0] == 0 and rp[-1] == npart and rp.size == nnode + 1
    assert np.array_equal(t["node"].numpy().T, cell_nodes)
    part = np.zeros(npart)
    for b in range(nb):
        ncell = min(B, nC - b * B)
        for q in range(bq[b], bq[b + 1]):
            sl = lsl[lrp[q]:lrp[q + 1]]
            assert np.all(sl >= 0) and np.all(sl < ncell * k)
            assert np.all(np.diff(sl) > 0)                      # slot order
            nodes = cell_nodes.reshape(-1)[b * B * k + sl]
            assert np.all(nodes == nodes[0])                    # one node per partial
            part[q] = vals.reshape(-1)[b * B * k + sl].sum()
    out = np.array([part[pe[rp[a]:rp[a + 1]]].sum() for a in range(nnode)])
    return out, t


@pytest.mark.parametrize("seed,nC,k,nnode,B", [(0, 1000, 10, 300, 256), (1, 513, 4, 97, 256),
                                               (2, 77, 10, 50, 16), (3, 256, 4, 40, 256)])
def test_block_partials_sum_like_a_scatter_add(seed, nC, k, nnode, B):
    rng = np.random.default_rng(seed)
    cell_nodes = np.stack([rng.choice(nnode, size=k, replace=False) for _ in range(nC)])
    vals = rng.standard_normal((nC, k))
    got, t = _emulate(cell_nodes, nnode, B, vals)
    ref = np.zeros(nnode)
    np.add.at(ref, cell_nodes.reshape(-1), vals.reshape(-1))
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_block_partials_of_a_cube_mesh():
    """The real P2 and P1 tables of a small cube: fewer partials than
    (cell, node) slots, and a node's partials from increasing blocks."""
    from waterscapes_amd.fem import MixedSpace
    from waterscapes_amd.mesh import UnitCubeMesh
    space = MixedSpace(UnitCubeMesh(6, 6, 6), 2)
    dm = space.dofmap
    for table, nnode in ((dm.cell_p2, dm.nP2), (dm.cell_p1, dm.nP1)):
        table = np.asarray(table, dtype=np.int64)
        vals = np.random.default_rng(5).standard_normal(table.shape)
        got, t = _emulate(table, nnode, 64, vals)
        ref = np.zeros(nnode)
        np.add.at(ref, table.reshape(-1), vals.reshape(-1))
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)
        assert t["npart"] < table.size
        blocks_of = np.empty(t["npart"], dtype=np.int64)
        bq = t["bq"].numpy()
        for b in range(bq.size - 1):
            blocks_of[t["gp"].numpy()[bq[b]:bq[b + 1]]] = b
        rp = t["rp"].numpy()
        for a in range(nnode):
            assert np.all(np.diff(blocks_of[rp[a]:rp[a + 1]]) > 0)
