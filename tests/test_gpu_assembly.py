"""GPU assembly / operator parity against the FP64 oracle.

Assembled matrices match the oracle to 1e-12 relative entrywise with an
identical CSR pattern (canonical numbering); operator applications match
scipy products of the oracle matrices."""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import CFG1, cfg2_params, lognormal_K, oracle_mesh
from oracle import mpet_oracle as O

pytestmark = pytest.mark.gpu

ENTRY_RTOL = 1e-12


def _build(n, params, K_cells=None, dt=0.01, theta=1.0):
    import torch
    from waterscapes_amd.fem import MixedSpace
    from waterscapes_amd.mesh import UnitCubeMesh
    from waterscapes_amd.device import DeviceSystem
    mesh = UnitCubeMesh(n, n, n)
    sp_ = MixedSpace(mesh, params["J"])
    dev = DeviceSystem(sp_, params, dt, theta, K_cells=K_cells, build_csr=True)
    torch.cuda.synchronize()
    return sp_, dev


def _csr_host(t):
    rp, ci, vals = t
    n = rp.numel() - 1
    return sp.csr_matrix((vals.cpu().numpy(), ci.cpu().numpy().astype(np.int64),
                          rp.cpu().numpy()), shape=(n, n))


def _entrywise(G, R):
    assert np.array_equal(G.indptr, R.indptr), "row pointers differ"
    assert np.array_equal(G.indices, R.indices), "column indices differ"
    scale = np.abs(R.data).max()
    err = np.abs(G.data - R.data)
    tol = ENTRY_RTOL * np.maximum(np.abs(R.data), 1e-3 * scale)
    bad = err > tol
    assert not bad.any(), f"{bad.sum()} entries off; worst rel {(err / np.maximum(np.abs(R.data), 1e-300)).max():.3e}"


@pytest.mark.parametrize("n,params,hetero", [(3, CFG1, False), (2, cfg2_params(4), True),
                                             (4, cfg2_params(3), True)])
def test_assembled_A_and_P_match_oracle(n, params, hetero):
    from waterscapes_amd.mesh import UnitCubeMesh
    nC = 6 * n ** 3
    K = lognormal_K(nC, params["K"], seed=3) if hetero else None
    space, dev = _build(n, params, K_cells=K, dt=0.0125, theta=1.0)
    m = oracle_mesh(space)
    A_ref = O.assemble_matrix(m["coords"], m["cells"], m["cell_p2"], m["cell_p1"], m["nP2"],
                              m["nP1"], params, 0.0125, 1.0, K_cells=K, which="A")
    P_ref = O.assemble_matrix(m["coords"], m["cells"], m["cell_p2"], m["cell_p1"], m["nP2"],
                              m["nP1"], params, 0.0125, 1.0, K_cells=K, which="P")
    A_ref.eliminate_zeros() if False else None
    _entrywise(_csr_host(dev.csr), A_ref)
    _entrywise(_csr_host(dev.expand_csr(1)), P_ref)


def test_operator_apply_matches_scipy():
    import torch
    params = cfg2_params(4)
    n = 3
    K = lognormal_K(6 * n ** 3, params["K"], seed=5)
    space, dev = _build(n, params, K_cells=K, dt=0.0125)
    A = _csr_host(dev.csr)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(dev.n)
    b = rng.standard_normal(dev.n)
    xt = torch.tensor(x, device="cuda")
    bt = torch.tensor(b, device="cuda")
    ref = A @ x
    for masked in (False,):
        y = dev.apply(xt, masked=masked).cpu().numpy()
        assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()
        y = dev.csr_apply(xt, masked=masked).cpu().numpy()
        assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()
        r = dev.apply(xt, b=bt, masked=masked).cpu().numpy()
        assert np.abs(r - (b - ref)).max() <= 1e-12 * np.abs(b - ref).max()
    # masked rows -> 0
    dofs = np.arange(0, dev.n, 7)
    dev.set_dirichlet(dofs)
    y = dev.apply(xt).cpu().numpy()
    ref_m = ref.copy()
    ref_m[dofs] = 0
    assert np.abs(y - ref_m).max() <= 1e-12 * np.abs(ref).max()
    # FP32 structured operator
    y32 = dev.apply(xt.float(), dtype=1, masked=False).double().cpu().numpy()
    assert np.linalg.norm(y32 - ref) <= 1e-5 * np.linalg.norm(ref)
    for lpr in (4, 8, 32, 64):
        y = dev.csr_apply(xt, lpr=lpr, masked=False).cpu().numpy()
        assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()


def test_rhs_pressure_matches_oracle():
    import torch
    params = CFG1
    n = 3
    space, dev = _build(n, params, dt=0.01, theta=0.5)
    m = oracle_mesh(space)
    prob = O.OracleProblem(m, params, dict(dt=0.01, theta=0.5, T=0.1))
    rng = np.random.default_rng(1)
    up = rng.standard_normal(dev.n)
    ref = prob.B_matrix() @ up
    b = torch.empty(dev.n, dtype=torch.float64, device="cuda")
    dev.rhs_pressure(torch.tensor(up, device="cuda"), b)
    got = b.cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_assembly_is_bitwise_reproducible():
    import torch
    params = cfg2_params(2)
    space, dev = _build(3, params, K_cells=lognormal_K(162, params["K"], seed=9))
    first = {k: v.clone() for k, v in dev.prim_t.items()}
    dev.assemble_primitives()
    torch.cuda.synchronize()
    for k, v in dev.prim_t.items():
        assert torch.equal(v, first[k]), k
