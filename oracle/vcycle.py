"""numpy restatement of the product's block preconditioner (AMG V-cycles
with Chebyshev smoothing on given hierarchies).

TEST INFRASTRUCTURE ONLY (see mpet_oracle.py header): the checker of the
device V-cycles (tests/test_gpu_solver.py, tests/test_batch_pack.py,
tests/test_gpu_batch.py) and the host side of bench.py's CPU baseline.

The preconditioner is the block diagonal of `prec`
(mpet/mpettotalpressuresolver.py:279-283) with Dirichlet rows and columns
eliminated (:445-446); each block is approximated by one V-cycle of a
smoothed-aggregation hierarchy (the reference uses hypre BoomerAMG, :448):
  * Chebyshev smoothing of D^-1 A on [rho_max / ratio, rho_max], rho_max =
    safety * rho, the recurrence of Saad, Iterative Methods for Sparse
    Linear Systems (2003), Alg. 12.1: step 0 d = r / theta, step i
    d = cd_i d + cr_i D^-1 (b - A x); pre-smoothing from x = 0,
    post-smoothing from the corrected x, the Dirichlet rows held at 0;
  * residual, restriction P^T, recursion, prolongation P;
  * a dense inverse on the coarsest level.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def chebyshev(lmin, lmax, degree):
    """(cd, cr) of the Chebyshev iteration on [lmin, lmax] (Saad Alg. 12.1)."""
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    sigma = theta / delta
    rho = 1.0 / sigma
    cd, cr = [0.0], [1.0 / theta]
    for _ in range(1, degree):
        rn = 1.0 / (2.0 * sigma - rho)
        cd.append(rn * rho)
        cr.append(2.0 * rn / delta)
        rho = rn
    return np.array(cd), np.array(cr)


def vcycle(levels, cinv, A0, mask0, prm, b):
    """One V-cycle for A0 x = b (columns of b solved at once).  levels: the
    hierarchy dicts (n, nc, A_rp/A_ci/A_v, P_rp/P_ci/P_v, dinv, rho); A0 the
    level-0 operator to use (None: the level's own matrix); mask0 the
    level-0 rows held at zero; prm: degree(l), eig_ratio, eig_safety."""
    degs = [prm.degree(l) for l in range(len(levels))]
    ops = [A0 if A0 is not None else sp.csr_matrix(
        (levels[0]["A_v"], levels[0]["A_ci"], levels[0]["A_rp"]), shape=(levels[0]["n"],) * 2)]
    ops += [sp.csr_matrix((lev["A_v"], lev["A_ci"], lev["A_rp"]), shape=(lev["n"],) * 2)
            for lev in levels[1:]]
    Ps = [sp.csr_matrix((lev["P_v"], lev["P_ci"], lev["P_rp"]), shape=(lev["n"], lev["nc"]))
          if lev["nc"] else None for lev in levels]

    def smooth(l, b, x=None):
        lev = levels[l]
        deg = degs[l]
        D = lev["dinv"][:, None]
        lmax = lev["rho"] * prm.eig_safety
        cd, cr = chebyshev(lmax / prm.eig_ratio, lmax, deg)
        m = mask0 if (l == 0 and mask0 is not None) else np.zeros(lev["n"], bool)
        if x is None:
            d = cr[0] * D * b
            d[m] = 0
            x = d.copy()
            start = 1
        else:
            d = np.zeros_like(b)
            start = 0
        for i in range(start, deg):
            d = cd[i] * d + cr[i] * D * (b - ops[l] @ x)
            d[m] = 0
            x = x + d
            x[m] = 0
        return x

    def vc(l, b):
        if l == len(levels) - 1:
            return cinv @ b
        m = mask0 if (l == 0 and mask0 is not None) else np.zeros(levels[l]["n"], bool)
        x = smooth(l, b)
        r_ = b - ops[l] @ x
        r_[m] = 0
        x = x + Ps[l] @ vc(l + 1, Ps[l].T @ r_)
        return smooth(l, b, x)

    return vc(0, b)


def block_preconditioner(solver, r, compact=False, A0=None):
    """z = P^-1 r with the data of a product KrylovSolver (its host
    hierarchies, masks and level-0 Laplacian): the displacement block
    (mu * vector Laplacian, three components at once) and the pressure
    blocks in field-major order.  compact: use the half-precision level-0
    copy of the FP32 hierarchy; A0: a host level-0 matrix to use instead
    of the device Laplacian."""
    s = solver.sys
    cfg = solver.cfg
    if A0 is None:
        rp, ci = s.pat_host["22"]
        lap = s.prim_t["Lap"].cpu().numpy()
        c = getattr(solver, "compact", None) if compact else None
        if c is not None:
            lap = c["A_v16"].float().cpu().numpy().astype(np.float64) * c["vscale"]
        A0 = sp.csr_matrix((lap, ci, rp), shape=(s.nP2, s.nP2))
    mask0 = solver.mask0.cpu().numpy().astype(bool)
    nu = s.nu
    zu = vcycle(solver.levels, solver.cinv, A0, mask0, cfg.amg,
                r[:nu].reshape(-1, 3) / s.mu).reshape(-1)
    F = s.F
    rp_fm = r[nu:].reshape(s.nP1, F).T.reshape(-1, 1)          # field-major
    zp = vcycle(solver.plevels, solver.pcinv, solver.Pp_host, None, cfg.pamg, rp_fm)
    zp = zp.reshape(F, s.nP1).T.reshape(-1)
    return np.concatenate([zu, zp])
