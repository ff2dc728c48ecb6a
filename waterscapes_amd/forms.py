"""Scalar coefficients of the total-pressure MPET forms and Dirichlet DOFs.

The discrete operator is a fixed linear combination of a few primitive
element tensors (see csrc/assemble.hip).  This module computes the
combination weights from the problem parameters, term by term from
mpet/mpettotalpressuresolver.py:262-283, and the Dirichlet DOF sets of
create_dirichlet_bcs (:123-144).
"""
from __future__ import annotations

import numpy as np

from . import refel


def lame(E, nu):
    """(mu, lambda) as in convert_to_mu_lmbda (mpet/mpetproblem.py:13-16)."""
    E = float(E)
    nu = float(nu)
    return E / (2.0 * (1.0 + nu)), nu * E / ((1.0 - 2.0 * nu) * (1.0 + nu))


def _s(S, i, j):
    return float(S[i][j])


def matrix_coefficients(alpha, c, S, lmbda, dt, theta):
    """Pressure block of lhs(F): C[f][g] multiplies the P1 mass, stiff[f] the
    K_f-weighted stiffness (f = 0 is p_0).

      (p0,w0) -1/lambda; (p_j,w0), (p0,w_i) -alpha/lambda        (:269, :271)
      (p_j,w_i) -(c_i delta_ij + alpha_i alpha_j / lambda)        (:270-271)
                -dt theta (delta_ij sum_k S_ik - S_ij)            (:273-274)
      stiffness -dt theta K_i                                     (:272)"""
    A = len(alpha)
    C = np.zeros((A + 1, A + 1))
    C[0, 0] = -1.0 / lmbda
    for j in range(A):
        C[0, j + 1] = C[j + 1, 0] = -float(alpha[j]) / lmbda
    for i in range(A):
        for j in range(A):
            v = -float(alpha[i]) * float(alpha[j]) / lmbda
            if i == j:
                v -= float(c[i])
                v -= dt * theta * sum(_s(S, i, k) for k in range(A) if k != i)
            else:
                v += dt * theta * _s(S, i, j)
            C[i + 1, j + 1] = v
    stiff = np.zeros(A + 1)
    stiff[1:] = -dt * theta
    return C, stiff


def rhs_coefficients(alpha, c, S, lmbda, dt, theta):
    """Part of rhs(F) linear in up_ (evaluated at t_n, :461):
      -c_i p_i_ - alpha_i/lambda (p0_ + sum_j alpha_j p_j_)
      + dt (1-theta) (K_i grad p_i_, grad w_i) + dt (1-theta) sum_j S_ij (p_i_ - p_j_)."""
    A = len(alpha)
    B = np.zeros((A + 1, A + 1))
    for i in range(A):
        B[i + 1, 0] = -float(alpha[i]) / lmbda
        for j in range(A):
            v = -float(alpha[i]) * float(alpha[j]) / lmbda
            if i == j:
                v -= float(c[i])
                v += dt * (1 - theta) * sum(_s(S, i, k) for k in range(A) if k != i)
            else:
                v -= dt * (1 - theta) * _s(S, i, j)
            B[i + 1, j + 1] = v
    stiff = np.zeros(A + 1)
    stiff[1:] = dt * (1 - theta)
    return B, stiff


def preconditioner_coefficients(alpha, c, S, lmbda, dt, theta):
    """Diagonal of the block preconditioner prec (:279-283):
      u: mu grad:grad; p0: mass; p_i: (alpha_i^2/lambda + c_i + dt theta
      sum_{j != i} S_ij) mass + dt theta K_i stiffness."""
    A = len(alpha)
    diag = np.zeros(A + 1)
    diag[0] = 1.0
    for i in range(A):
        diag[i + 1] = (float(alpha[i]) ** 2 / lmbda + float(c[i])
                       + sum(dt * theta * _s(S, i, j) for j in range(A) if j != i))
    stiff = np.zeros(A + 1)
    stiff[1:] = dt * theta
    return diag, stiff


# ---------------------------------------------------------------------------
# two-field formulation (MPETSolver, mpet/mpetsolver.py:196-201, 249-254, 269-271)
# ---------------------------------------------------------------------------


def two_field_matrix_coefficients(c, S, dt, theta):
    """Pressure block of lhs(F) of the two-field form: C[i][j] multiplies the
    P1 mass, stiff[i] the K_i-weighted stiffness (fields p_1..p_A only):

      (p_j,q_i) -c_i delta_ij                                 (:198)
                -dt theta (delta_ij sum_k S_ik - S_ij)        (:201)
      stiffness -dt theta K_i                                 (:200)"""
    A = len(c)
    C = np.zeros((A, A))
    for i in range(A):
        for j in range(A):
            if i == j:
                C[i, j] = -float(c[i]) - dt * theta * sum(_s(S, i, k) for k in range(A) if k != i)
            else:
                C[i, j] = dt * theta * _s(S, i, j)
    return C, np.full(A, -dt * theta)


def two_field_rhs_coefficients(c, S, dt, theta):
    """Part of rhs(F) of the two-field form linear in the pressures of up_
    (at t_n): -c_i p_i_ + dt (1-theta) (K_i grad p_i_, grad q_i)
    + dt (1-theta) sum_j S_ij (p_i_ - p_j_).  The displacement part,
    -alpha_i (div u_, q_i) (:199), is added by the caller."""
    A = len(c)
    B = np.zeros((A, A))
    for i in range(A):
        for j in range(A):
            if i == j:
                B[i, j] = -float(c[i]) + dt * (1 - theta) * sum(_s(S, i, k)
                                                                 for k in range(A) if k != i)
            else:
                B[i, j] = -dt * (1 - theta) * _s(S, i, j)
    return B, np.full(A, dt * (1 - theta))


def two_field_preconditioner_coefficients(c, S, dt, theta):
    """Diagonal blocks of the two-field prec (:269-271): u mu grad:grad;
    p_i (c_i + dt theta sum_{j != i} S_ij) mass + dt theta K_i stiffness.
    The reference writes the stiffness weight as ``dt*theta**inner(...)``
    (theta raised to the power of the integrand, :270), evidently meaning
    dt*theta*inner(...); the product form is used here (its iterative path,
    the only user of prec, is switched off in the reference, :308-314)."""
    A = len(c)
    diag = np.array([float(c[i]) + sum(dt * theta * _s(S, i, j) for j in range(A) if j != i)
                     for i in range(A)])
    return diag, np.full(A, dt * theta)


def dirichlet_nodes(space, markers, field):
    """Nodes in the closure of exterior facets marked 0 (DOLFIN's
    topological DirichletBC): P2 nodes (3 vertices + 3 edges) for field 0,
    P1 nodes (3 vertices) otherwise."""
    mesh = space.mesh
    dm = space.dofmap
    vals = np.asarray(markers.array() if hasattr(markers, "array") else markers)
    sel = np.nonzero(vals == 0)[0]
    if sel.size == 0:
        return np.zeros(0, dtype=np.int64)
    cl = mesh.facet_cell[sel]
    lf = mesh.facet_local[sel]
    if field == 0:
        loc = np.array([refel.face_local_p2_nodes(f) for f in range(4)])[lf]   # nF,6
        nodes = dm.cell_p2[cl[:, None], loc]
    else:
        loc = np.array(refel.FACES)[lf]
        nodes = dm.cell_p1[cl[:, None], loc]
    return np.unique(nodes.reshape(-1)).astype(np.int64)


def field_dofs(space, field, nodes):
    """DOFs of sub-space `field` (0 = u, 1.. = the pressure fields in the
    space's order) at the given nodes."""
    if field == 0:
        return (3 * nodes[:, None] + np.arange(3)).reshape(-1)
    return space.nu + space.F * nodes + (field - 1)
