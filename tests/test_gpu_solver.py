"""GPU MINRES / preconditioner parity.

* The device preconditioner equals a numpy restatement of the same block
  preconditioner (V-cycle with Chebyshev smoothing on the same hierarchy,
  Chebyshev-Jacobi pressure blocks) to 1e-12 (FP64) / 1e-5 (FP32).
* One time-step solve (FP64 MINRES, and FP32 MINRES in FP64 refinement)
  matches the oracle's direct solve of the symmetrically eliminated system
  (mpet/bc_symmetric.py:11-18) to the stated relative error."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import cfg2_params, lognormal_K, oracle_mesh
from oracle import mpet_oracle as O

pytestmark = pytest.mark.gpu


def _setup(n=4, A=4, mode="fp64", seed=0, **kw):
    import torch
    from waterscapes_amd.fem import MixedSpace
    from waterscapes_amd.mesh import UnitCubeMesh, CompiledSubDomain, MeshFunction
    from waterscapes_amd.device import DeviceSystem
    from waterscapes_amd.krylov import KrylovSolver, SolverConfig
    from waterscapes_amd import forms
    params = cfg2_params(A)
    mesh = UnitCubeMesh(n, n, n)
    space = MixedSpace(mesh, A)
    K = lognormal_K(mesh.num_cells(), params["K"], seed=seed)
    dev = DeviceSystem(space, params, 0.0125, 1.0, K_cells=K)
    mm = MeshFunction("size_t", mesh, 2, 2**63 - 1)
    CompiledSubDomain("on_boundary && near(x[2], 0.0)").mark(mm, 0)
    cm = MeshFunction("size_t", mesh, 2, 2**63 - 1)
    CompiledSubDomain("on_boundary && near(x[2], 1.0)").mark(cm, 0)
    dofs = [forms.field_dofs(space, 0, forms.dirichlet_nodes(space, mm, 0))]
    for i in range(A):
        dofs.append(forms.field_dofs(space, i + 2, forms.dirichlet_nodes(space, cm, i + 2)))
    cfg = SolverConfig(mode=mode, **kw)
    solver = KrylovSolver(dev, dofs, cfg)
    torch.cuda.synchronize()
    return params, space, dev, solver, dofs, K


def _ref_precond(solver, r, dtype=np.float64):
    """numpy restatement of the block preconditioner on the same data."""
    from waterscapes_amd import amg
    s = solver.sys
    cfg = solver.cfg
    levels = solver.levels
    deg = cfg.amg.smoother_degree
    ops, Ps = [], []
    for l, lev in enumerate(levels):
        if l == 0:
            rp, ci = s.pat_host["22"]
            A0 = sp.csr_matrix((s.prim_t["Lap"].cpu().numpy(), ci, rp), shape=(s.nP2, s.nP2))
            ops.append(A0)
        else:
            ops.append(sp.csr_matrix((lev["A_v"], lev["A_ci"], lev["A_rp"]), shape=(lev["n"],) * 2))
        Ps.append(sp.csr_matrix((lev["P_v"], lev["P_ci"], lev["P_rp"]), shape=(lev["n"], lev["nc"]))
                  if lev["nc"] else None)
    mask0 = solver.mask0.cpu().numpy().astype(bool)

    def smooth(l, b, x=None):
        lev = levels[l]
        A = ops[l]
        D = lev["dinv"][:, None]
        lmax = lev["rho"] * cfg.amg.eig_safety
        cd, cr = amg.chebyshev_coefficients(lmax / cfg.amg.eig_ratio, lmax, deg)
        m = mask0 if l == 0 else np.zeros(lev["n"], bool)
        if x is None:
            d = cr[0] * D * b
            d[m] = 0
            x = d.copy()
            start = 1
        else:
            d = np.zeros_like(b)
            start = 0
        for i in range(start, deg):
            res = b - A @ x
            d = cd[i] * d + cr[i] * D * res
            d[m] = 0
            x = x + d
            x[m] = 0
        return x

    def vc(l, b):
        if l == len(levels) - 1:
            return solver.cinv @ b
        m = mask0 if l == 0 else np.zeros(levels[l]["n"], bool)
        x = smooth(l, b)
        r_ = b - ops[l] @ x
        r_[m] = 0
        e = vc(l + 1, Ps[l].T @ r_)
        x = x + Ps[l] @ e
        return smooth(l, b, x)

    nu = s.nu
    zu = vc(0, r[:nu].reshape(-1, 3) / s.mu).reshape(-1)
    # pressure blocks
    F = s.F
    rp, ci = s.pat_host["11"]
    M1 = sp.csr_matrix((s.prim_t["M1"].cpu().numpy(), ci, rp), shape=(s.nP1,) * 2)
    St = s.prim_t["St"].cpu().numpy().reshape(s.A, -1)
    rpz = r[nu:].reshape(s.nP1, F)
    zp = np.zeros_like(rpz)
    pmask = solver.pmask.cpu().numpy().reshape(s.nP1, F).astype(bool)
    dinvp = solver.dinvp_host.reshape(s.nP1, F)
    for f in range(F):
        Pf = solver.Pd[f] * M1
        if f >= 1:
            Pf = Pf + s.Ps[f] * sp.csr_matrix((St[f - 1], ci, rp), shape=(s.nP1,) * 2)
        cd, cr = amg.chebyshev_coefficients(solver.plmin[f], solver.plmax[f], solver.pdeg)
        b = rpz[:, f]
        m = pmask[:, f]
        d = cr[0] * dinvp[:, f] * b
        d[m] = 0
        x = d.copy()
        for i in range(1, solver.pdeg):
            d = cd[i] * d + cr[i] * dinvp[:, f] * (b - Pf @ x)
            d[m] = 0
            x = x + d
            x[m] = 0
        zp[:, f] = x
    return np.concatenate([zu, zp.reshape(-1)])


def test_precond_matches_numpy_restatement():
    import torch
    params, space, dev, solver, dofs, K = _setup(n=4, A=3)
    rng = np.random.default_rng(2)
    r = rng.standard_normal(dev.n)
    r[np.concatenate(dofs)] = 0.0
    z = solver.precond(torch.tensor(r, device="cuda"), 0).cpu().numpy()
    ref = _ref_precond(solver, r)
    assert np.abs(z - ref).max() <= 1e-11 * np.abs(ref).max()
    assert np.all(z[np.concatenate(dofs)] == 0.0)


def test_precond_fp32_close_to_fp64():
    import torch
    params, space, dev, solver, dofs, K = _setup(n=4, A=2, mode="mixed")
    rng = np.random.default_rng(3)
    r = rng.standard_normal(dev.n)
    r[np.concatenate(dofs)] = 0.0
    z = solver.precond(torch.tensor(r, device="cuda", dtype=torch.float32), 1).double().cpu().numpy()
    ref = _ref_precond(solver, r)
    assert np.linalg.norm(z - ref) <= 1e-5 * np.linalg.norm(ref)


def _oracle_step(space, params, K, dofs, up_prev, g):
    m = oracle_mesh(space)
    A = O.assemble_matrix(m["coords"], m["cells"], m["cell_p2"], m["cell_p1"], m["nP2"],
                          m["nP1"], params, 0.0125, 1.0, K_cells=K, which="A")
    prob = O.OracleProblem(m, params, dict(dt=0.0125, theta=1.0, T=1.0), K_cells=K)
    b = prob.B_matrix() @ up_prev
    for d in dofs:
        b[d] = g[d]
        A = O.apply_symmetric(A, d, b)
    return spla.spsolve(A.tocsc(), b), b


@pytest.mark.parametrize("mode,tol", [("fp64", 1e-8), ("mixed", 1e-8)])
def test_step_solve_matches_oracle_direct(mode, tol):
    import torch
    params, space, dev, solver, dofs, K = _setup(n=4, A=4, mode=mode)
    rng = np.random.default_rng(11)
    up_prev = rng.standard_normal(dev.n)
    g = np.zeros(dev.n)
    alld = np.concatenate(dofs)
    g[alld] = rng.standard_normal(alld.size)
    ref, b_el = _oracle_step(space, params, K, dofs, up_prev, g)
    up_t = torch.tensor(up_prev, device="cuda")
    b = torch.empty_like(up_t)
    dev.rhs_pressure(up_t, b)
    x = up_t.clone()
    x[torch.as_tensor(alld, device="cuda")] = torch.tensor(g[alld], device="cuda")
    st = solver.solve(b, x, bnorm=float(np.linalg.norm(b_el)))
    got = x.cpu().numpy()
    # per-field relative L2 error
    for f, idx in enumerate([space.field_index(i) for i in range(space.A + 2)]):
        e = np.linalg.norm(got[idx] - ref[idx]) / max(np.linalg.norm(ref[idx]), 1e-300)
        assert e <= tol, (mode, f, e, st.inner_iterations)
    assert st.converged
