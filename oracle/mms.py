"""Manufactured solutions for the total-pressure MPET equations (3D).

TEST INFRASTRUCTURE ONLY (see mpet_oracle.py header).

Restates the construction of test_convergence_totalpressuresolver.py:20-101
in three dimensions: p_0 = lambda div u - sum_j alpha_j p_j (:55),
f = -(div 2 mu eps(u) + grad p_0) (:80) and the NEGATED source convention for
g_i (:86-89; also demo_adaptive.py:48-49):
  g_i = -c_i dp_i/dt - alpha_i/lambda d/dt (p_0 + sum_j alpha_j p_j)
        + div(K_i grad p_i) - sum_j S_ij (p_i - p_j).
The default solution is the cfg1 workload's: u_k = 0.1 t sin(2 pi x) sin(2 pi y)
sin(2 pi z) for every component k and p_j = (j+1) t cos(pi x) cos(pi y) cos(pi z).
"""
from __future__ import annotations

import numpy as np
import sympy as S

X, Y, Z, T = S.symbols("x0 x1 x2 t")


def exact_solutions(params, u_expr=None, p_exprs=None):
    """Return sympy expressions (u[3], p[A+1] incl. p_0, f[3], g[A])."""
    A = int(params["J"])
    E, nu = params["E"], params["nu"]
    alpha, c, K, Smat = params["alpha"], params["c"], params["K"], params["S"]
    lmbda = nu * E / ((1.0 - 2.0 * nu) * (1.0 + nu))
    mu = E / (2.0 * (1.0 + nu))
    xs = (X, Y, Z)
    if u_expr is None:
        base = 0.1 * T * S.sin(2 * S.pi * X) * S.sin(2 * S.pi * Y) * S.sin(2 * S.pi * Z)
        u = [base, base, base]
    else:
        u = list(u_expr)
    if p_exprs is None:
        pb = T * S.cos(S.pi * X) * S.cos(S.pi * Y) * S.cos(S.pi * Z)
        p = [None] + [(j + 1) * pb for j in range(A)]
    else:
        p = [None] + list(p_exprs)
    div_u = sum(S.diff(u[i], xs[i]) for i in range(3))
    p[0] = lmbda * div_u - sum(alpha[j] * p[j + 1] for j in range(A))
    grad_u = [[S.diff(u[i], xs[j]) for j in range(3)] for i in range(3)]
    eps = [[0.5 * (grad_u[i][j] + grad_u[j][i]) for j in range(3)] for i in range(3)]
    sig = [[2 * mu * eps[i][j] for j in range(3)] for i in range(3)]
    f = [-(sum(S.diff(sig[i][j], xs[j]) for j in range(3)) + S.diff(p[0], xs[i]))
         for i in range(3)]
    g = []
    for i in range(A):
        gi = (-c[i] * S.diff(p[i + 1], T)
              - alpha[i] / lmbda * S.diff(p[0] + sum(alpha[j] * p[j + 1] for j in range(A)), T)
              + sum(S.diff(K[i] * S.diff(p[i + 1], xs[j]), xs[j]) for j in range(3))
              - sum(Smat[i][j] * (p[i + 1] - p[j + 1]) for j in range(A)))
        g.append(gi)
    return u, p, f, g


def boundary_data(params, u, p, beta):
    """Natural boundary data of the exact solution (sympy):
      sigma = 2 mu eps(u) + p_0 I, traction s = sigma . n (the reference
      test's sigma, test_convergence_totalpressuresolver.py:60-70, :133-138);
      network flux I_i = -K_i grad p_i . n (the weak form's boundary term
      +dt int I_i w_i, :322, against -dt int (K_i grad p_i . n) w_i from
      integrating div(K_i grad p_i) by parts);
      Robin data p_robin_i = p_i + (K_i grad p_i . n) / beta_i, so that
      K_i grad p_i . n = -beta_i (p_i - p_robin_i) (:324).
    Returns (sigma 3x3, grad_p list of 3-vectors per network)."""
    E, nu = params["E"], params["nu"]
    mu = E / (2.0 * (1.0 + nu))
    xs = (X, Y, Z)
    grad_u = [[S.diff(u[i], xs[j]) for j in range(3)] for i in range(3)]
    sigma = [[mu * (grad_u[i][j] + grad_u[j][i]) + (p[0] if i == j else 0) for j in range(3)]
             for i in range(3)]
    grads = [[S.diff(p[i + 1], v) for v in xs] for i in range(int(params["J"]))]
    return sigma, grads


def lambdify(expr):
    """numpy callable f(X (n,3), t) -> (n,) for a scalar sympy expression."""
    fn = S.lambdify((X, Y, Z, T), expr, modules="numpy")

    def call(P, t):
        P = np.asarray(P, dtype=np.float64)
        return np.broadcast_to(np.asarray(fn(P[:, 0], P[:, 1], P[:, 2], t), float),
                               (P.shape[0],)).copy()
    return call


def lambdify_vec(exprs):
    fns = [lambdify(e) for e in exprs]

    def call(P, t):
        return np.stack([f(P, t) for f in fns], axis=1)
    return call


def grad_lambdify(expr):
    return [lambdify(S.diff(expr, v)) for v in (X, Y, Z)]


def error_norms(coords, cells, cell_nodes, x_nodes, values, exact, t, degree, order=8):
    """L2 and H1-seminorm errors of a P1 (degree 1, cell_nodes = vertices) or
    P2 (degree 2, cell_nodes = 10 UFC nodes) scalar field against the sympy
    expression `exact`, by a degree-`order` quadrature per cell."""
    from .mpet_oracle import tet_rule, cell_geometry, p2_ref_values, p2_ref_grads, \
        p1_ref_values, P1_REF_GRADS
    J, det, Jinv = cell_geometry(coords, cells)
    qp, qw = tet_rule(order)
    v0 = coords[cells[:, 0]]
    Xq = v0[:, None, :] + np.einsum("cij,qj->cqi", J, qp)
    if degree == 2:
        phi = p2_ref_values(qp)
        dphi = np.einsum("ckm,qak->cqam", Jinv, p2_ref_grads(qp))
    else:
        phi = p1_ref_values(qp)
        g1 = np.einsum("ckm,ak->cam", Jinv, P1_REF_GRADS)
        dphi = np.broadcast_to(g1[:, None], (cells.shape[0], len(qp), 4, 3))
    coef = values[cell_nodes]
    uh = np.einsum("qa,ca->cq", phi, coef)
    duh = np.einsum("cqam,ca->cqm", dphi, coef)
    ue = lambdify(exact)(Xq.reshape(-1, 3), t).reshape(uh.shape)
    gre = np.stack([g(Xq.reshape(-1, 3), t).reshape(uh.shape) for g in grad_lambdify(exact)], -1)
    l2 = np.sqrt(np.einsum("cq,q,c->", (ue - uh) ** 2, qw, det))
    h1 = np.sqrt(np.einsum("cq,q,c->", ((gre - duh) ** 2).sum(-1), qw, det))
    return l2, h1
