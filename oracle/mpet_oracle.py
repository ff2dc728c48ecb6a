"""FP64 CPU oracle for the total-pressure MPET time step.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (waterscapes_amd/) may
import this module; only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, as the checker.

It restates, with numpy/scipy and Gauss quadrature, the forms and the time /
boundary-condition logic of the reference:

  * bilinear form F, preconditioner prec, right-hand sides L0, L1, L2
    .................................. mpet/mpettotalpressuresolver.py:239-324
  * Dirichlet DOFs (closure of facets marked 0, u on sub(0), p_i on
    sub(i+2), no BC on p_0) ............ mpet/mpettotalpressuresolver.py:123-144
  * row replacement (DirichletBC.apply) for the direct path
    .................................. mpet/mpettotalpressuresolver.py:361-362,399-400
  * symmetric elimination (MatZeroRowsColumns, diag 1, lifting into b)
    .................................. mpet/bc_symmetric.py:6-22, called at :445-446,487-489
  * time loop t_theta = t + theta*dt, t_{n+1} = t_theta + (1-theta)*dt,
    while t < T - 1e-9 ................ mpet/mpettotalpressuresolver.py:369-411,456-502
  * direct LU (MUMPS in the reference; SuperLU here) .................... :365,403
  * preconditioned MINRES with the block-diagonal preconditioner ......... :448,493

It also restates the two-field MPETSolver (mpet/mpetsolver.py): forms
:196-201 with the stress of mpetproblem.py:18-24, Robin :252-254, prec
:269-271, Dirichlet DOFs on sub(i+1) (:74-81), and its direct time loop
:382-462 (OracleProblem(formulation="two_field")), and the nullspace
multipliers of both solvers (rigid motions of mpet/rm_basis_L2.py:10-74,
constraint terms mpettotalpressuresolver.py:285-303, mpetsolver.py:203-215)
as a bordered system solved by LU (OracleProblem(u_nullspace=...,
p_nullspace=...)); and the a-posteriori error indicators of
AdaptiveMPETSolver (mpet/adaptivempetsolver.py:82-185, error_indicators).

Canonical numbering (inputs to every function): u DOF = 3*a + j for P2 node a,
component j; p_f DOF = 3*nP2 + F*c + f for P1 node c, field f = 0..F-1 with
F = A+1 (p_0, p_1..p_A) for the total-pressure and F = A (p_1..p_A) for the
two-field formulation.

Pinning (see tests/test_oracle.py): element tensors against sympy-exact
fixtures (tests/golden/), P1 closed forms, kernels (rigid motions, constants),
symmetric-elimination == row-replacement solutions, 3D manufactured-solution
convergence rates at the reference's thresholds, and the uniform-pressure
known answer p_0 = -sum_j alpha_j P_j.
"""
from __future__ import annotations

import math
from itertools import product

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla
from scipy.special import roots_jacobi

EDGES = ((2, 3), (1, 3), (1, 2), (0, 3), (0, 2), (0, 1))
FACES = ((1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2))

# ---------------------------------------------------------------------------
# quadrature and Lagrange bases (own implementation: Vandermonde inversion)
# ---------------------------------------------------------------------------


def tet_rule(order):
    """Stroud conical-product rule on the unit tetrahedron (volume 1/6)."""
    n = order // 2 + 1
    a0, w0 = roots_jacobi(n, 0, 0)
    a1, w1 = roots_jacobi(n, 1, 0)
    a2, w2 = roots_jacobi(n, 2, 0)
    P, W = [], []
    for i, j, k in product(range(n), range(n), range(n)):
        z = (1 + a2[k]) / 2
        y = (1 + a1[j]) / 2 * (1 - z)
        x = (1 + a0[i]) / 2 * (1 - y - z)
        P.append((x, y, z))
        W.append(w0[i] * w1[j] * w2[k] / 64.0)
    return np.array(P), np.array(W)


def tri_rule(order):
    """Collapsed Gauss rule on the unit triangle (area 1/2)."""
    n = order // 2 + 1
    a0, w0 = roots_jacobi(n, 0, 0)
    a1, w1 = roots_jacobi(n, 1, 0)
    P, W = [], []
    for i, j in product(range(n), range(n)):
        y = (1 + a1[j]) / 2
        x = (1 + a0[i]) / 2 * (1 - y)
        P.append((x, y))
        W.append(w0[i] * w1[j] / 8.0)
    return np.array(P), np.array(W)


def _monomials(k, d):
    return [e for e in product(range(k + 1), repeat=d) if sum(e) <= k]


def lagrange_nodes(k, d=3):
    """P_k nodes in reference coordinates: barycentric multi-indices in the
    order vertices first, then lexicographic (same order as the product's
    Expression interpolation)."""
    n = d + 1
    if k == 0:
        return np.full((1, d), 1.0 / n)
    allidx = [a for a in product(range(k + 1), repeat=n) if sum(a) == k]
    vert = [tuple(k if s == i else 0 for s in range(n)) for i in range(n)]
    idx = vert + sorted(a for a in allidx if a not in vert)
    bary = np.array(idx, dtype=float) / k
    return bary[:, 1:]                       # reference coords = lambda_1..d


def lagrange_eval(k, pts, d=3):
    """Values (npts, nbasis) of the P_k Lagrange basis at reference points."""
    nodes = lagrange_nodes(k, d)
    mons = _monomials(k, d)
    V = np.array([[np.prod(p ** np.array(e)) for e in mons] for p in nodes])
    Pm = np.array([[np.prod(p ** np.array(e)) for e in mons] for p in pts])
    return Pm @ np.linalg.inv(V)


def p2_ref_values(pts):
    x, y, z = pts[:, 0], pts[:, 1], pts[:, 2]
    lam = np.stack([1 - x - y - z, x, y, z], axis=1)
    vals = [lam[:, i] * (2 * lam[:, i] - 1) for i in range(4)]
    vals += [4 * lam[:, i] * lam[:, j] for (i, j) in EDGES]
    return np.stack(vals, axis=1)


def p2_ref_grads(pts):
    """(npts, 10, 3) reference gradients of the P2 basis."""
    x, y, z = pts[:, 0], pts[:, 1], pts[:, 2]
    lam = np.stack([1 - x - y - z, x, y, z], axis=1)
    dlam = np.array([[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    g = [(4 * lam[:, i, None] - 1) * dlam[i] for i in range(4)]
    g += [4 * (lam[:, i, None] * dlam[j] + lam[:, j, None] * dlam[i]) for (i, j) in EDGES]
    return np.stack(g, axis=1)


def p1_ref_values(pts):
    x, y, z = pts[:, 0], pts[:, 1], pts[:, 2]
    return np.stack([1 - x - y - z, x, y, z], axis=1)


P1_REF_GRADS = np.array([[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)

# ---------------------------------------------------------------------------
# geometry
# ---------------------------------------------------------------------------


def cell_geometry(coords, cells):
    """Jacobian J (nC,3,3) with columns v_k - v_0, |det J|, J^{-1}."""
    v = coords[cells]                                     # nC,4,3
    J = np.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]], axis=2)
    det = np.linalg.det(J)
    Jinv = np.linalg.inv(J)
    return J, np.abs(det), Jinv


# ---------------------------------------------------------------------------
# element tensors
# ---------------------------------------------------------------------------


def element_tensors(coords, cells):
    """Per-cell FP64 element tensors by degree-2 quadrature (exact for the
    P2/P1 forms on affine cells).

    Returns dict with
      'E'   (nC,30,30)  int eps(phi):eps(psi) (unscaled; times 2 mu in F),
                        local DOF 3*a + j;
      'G'   (nC,10,10)  int grad phi_a . grad phi_b  (P2 Laplacian);
      'B'   (nC,30,4)   int psi_c div(phi_a e_j)   (row 3a+j, col c);
      'M'   (nC,4,4)    P1 mass;
      'S'   (nC,4,4)    P1 stiffness;
      'D'   (nC,30,30)  int div(phi_a e_i) div(phi_b e_j) (two-field lambda term);
    """
    _, det, Jinv = cell_geometry(coords, cells)
    qp, qw = tet_rule(2)
    rg = p2_ref_grads(qp)                                  # nq,10,3
    # physical gradient: grad phi = J^{-T} ref_grad
    pg = np.einsum("ckm,qak->cqam", Jinv, rg)              # nC,nq,10,3
    wdet = det[:, None] * qw[None, :]                      # nC,nq
    nC = cells.shape[0]
    # L2 product of gradients: Gm[c,a,b,m,n] = int d_m phi_a d_n phi_b
    Gm = np.einsum("cq,cqam,cqbn->cabmn", wdet, pg, pg)
    G = np.einsum("cabmm->cab", Gm)
    # eps(phi_a e_i):eps(phi_b e_j) = 1/2 (delta_ij grad.grad + d_j phi_a d_i phi_b)
    E = np.zeros((nC, 10, 3, 10, 3))
    for i in range(3):
        for j in range(3):
            E[:, :, i, :, j] = 0.5 * ((i == j) * G + Gm[:, :, :, j, i])
    E = E.reshape(nC, 30, 30)
    psi = p1_ref_values(qp)                                # nq,4
    Bt = np.einsum("cq,cqaj,qd->cajd", wdet, pg, psi)      # nC,10,3,4
    B = Bt.reshape(nC, 30, 4)
    M = np.einsum("cq,qa,qb->cab", wdet, psi, psi)
    g1 = np.einsum("ckm,ak->cam", Jinv, P1_REF_GRADS)      # nC,4,3
    S = (det / 6.0)[:, None, None] * np.einsum("cam,cbm->cab", g1, g1)
    # div(phi_b e_j) div(phi_a e_i) = d_i phi_a d_j phi_b  (row 3a+i, col 3b+j)
    D = np.transpose(Gm, (0, 1, 3, 2, 4)).reshape(nC, 30, 30)
    return {"E": E, "G": G, "B": B, "M": M, "S": S, "D": D, "det": det}


# ---------------------------------------------------------------------------
# parameters
# ---------------------------------------------------------------------------


def lame(E, nu):
    """mu, lambda (mpet/mpetproblem.py:13-16, mpettotalpressuresolver.py:262-263)."""
    mu = E / (2.0 * (1.0 + nu))
    lmbda = nu * E / ((1.0 - 2.0 * nu) * (1.0 + nu))
    return mu, lmbda


def pressure_coefficients(alpha, c, S, lmbda, dt, theta):
    """(A+1)x(A+1) coefficient matrices multiplying the P1 mass in the
    pressure rows: Cm for the matrix (lhs of F) and Bm for the up_ part of
    the right-hand side (rhs of F), mpettotalpressuresolver.py:269-274."""
    A = len(alpha)
    al = np.asarray(alpha, float)
    Cm = np.zeros((A + 1, A + 1))
    Bm = np.zeros((A + 1, A + 1))
    Cm[0, 0] = -1.0 / lmbda
    Cm[0, 1:] = -al / lmbda
    Cm[1:, 0] = -al / lmbda
    Bm[1:, 0] = -al / lmbda
    for i in range(A):
        for j in range(A):
            base = -(c[i] * (i == j) + al[i] * al[j] / lmbda)
            if i == j:
                tr = sum(S[i][k] for k in range(A) if k != i)
                Cm[i + 1, j + 1] = base - dt * theta * tr
                Bm[i + 1, j + 1] = base + dt * (1 - theta) * tr
            else:
                Cm[i + 1, j + 1] = base + dt * theta * S[i][j]
                Bm[i + 1, j + 1] = base - dt * (1 - theta) * S[i][j]
    return Cm, Bm


# ---------------------------------------------------------------------------
# global assembly
# ---------------------------------------------------------------------------


class Layout:
    """Canonical numbering with F pressure fields per P1 node: F = A + 1
    (total pressure: p_0, p_1..p_A) or F = A (two-field: p_1..p_A)."""

    def __init__(self, nP2, nP1, A, F=None):
        self.nP2, self.nP1, self.A = nP2, nP1, A
        self.F = A + 1 if F is None else F
        self.nu = 3 * nP2
        self.n = self.nu + self.F * nP1

    def net(self, i):
        """Layout field of network i (0-based)."""
        return i + (self.F - self.A)

    def udofs(self, cell_p2):
        """(nC,30) local u DOFs 3a+j."""
        return (3 * cell_p2[:, :, None] + np.arange(3)).reshape(cell_p2.shape[0], 30)

    def pdofs(self, cell_p1, f):
        return self.nu + self.F * cell_p1 + f


def _coo(rows, cols, vals, n):
    return sp.coo_matrix((vals.reshape(-1), (rows.reshape(-1), cols.reshape(-1))),
                         shape=(n, n))


def _blk(ld, rd, val):
    """Broadcast local row dofs (nC,r) / col dofs (nC,c) against val (nC,r,c)."""
    R = np.broadcast_to(ld[:, :, None], val.shape)
    Cc = np.broadcast_to(rd[:, None, :], val.shape)
    return R, Cc


def assemble_matrix(coords, cells, cell_p2, cell_p1, nP2, nP1, params,
                    dt, theta, K_cells=None, robin=None, which="A"):
    """Global FP64 matrix in CSR (duplicates summed, sorted indices).

    which == "A": lhs(F) + sum_i lhs(L2_i) (mpettotalpressuresolver.py:267-274,
    324; assembled at :437-441).  which == "P": the preconditioner form prec
    (:279-283), no boundary conditions.  Pattern: every block the form couples
    (u-u, u-p0, p0-u, all p-p pairs); u-p_i (i >= 1) is structurally empty.

    K_cells: (nC, A) per-cell K_i (DG0) or None to use params['K'].
    robin: optional list over networks of (facet_vertex_triples, beta_values
           at facet nodes) – see robin_facet_mass.
    """
    A = int(params["J"])
    mu, lmbda = lame(params["E"], params["nu"])
    alpha, c, S = params["alpha"], params["c"], params["S"]
    lay = Layout(nP2, nP1, A)
    et = element_tensors(coords, cells)
    nC = cells.shape[0]
    if K_cells is None:
        K_cells = np.tile(np.asarray(params["K"], float), (nC, 1))
    ud = lay.udofs(cell_p2)
    blocks = []
    if which == "A":
        val = 2 * mu * et["E"]
        blocks.append((ud, ud, val))
        pd0 = lay.pdofs(cell_p1, 0)
        blocks.append((ud, pd0, et["B"]))                          # p0 div v
        blocks.append((pd0, ud, np.transpose(et["B"], (0, 2, 1))))  # div u w0
        Cm, _ = pressure_coefficients(alpha, c, S, lmbda, dt, theta)
        for f in range(A + 1):
            for g in range(A + 1):
                v = Cm[f, g] * et["M"]
                if f == g and f >= 1:
                    v = v - dt * theta * K_cells[:, f - 1, None, None] * et["S"]
                blocks.append((lay.pdofs(cell_p1, f), lay.pdofs(cell_p1, g), v))
    else:
        G = mu * et["G"]
        Gu = np.zeros((nC, 10, 3, 10, 3))
        for j in range(3):
            Gu[:, :, j, :, j] = G
        blocks.append((ud, ud, Gu.reshape(nC, 30, 30)))
        blocks.append((lay.pdofs(cell_p1, 0), lay.pdofs(cell_p1, 0), et["M"]))
        for i in range(A):
            a_i = alpha[i] ** 2 / lmbda + c[i] + sum(dt * theta * S[i][j]
                                                     for j in range(A) if j != i)
            v = a_i * et["M"] + dt * theta * K_cells[:, i, None, None] * et["S"]
            pdi = lay.pdofs(cell_p1, i + 1)
            blocks.append((pdi, pdi, v))
    rows, cols, vals = [], [], []
    for (ld, rd, v) in blocks:
        R, Cc = _blk(ld, rd, v)
        rows.append(R.reshape(-1))
        cols.append(Cc.reshape(-1))
        vals.append(v.reshape(-1))
    if which == "A" and robin is not None:
        for i, rb in enumerate(robin):
            if rb is None:
                continue
            tri_dofs, Mf = rb                        # (nF,3) P1 nodes, (nF,3,3)
            pdi = lay.nu + (A + 1) * tri_dofs + (i + 1)
            R, Cc = _blk(pdi, pdi, Mf)
            rows.append(R.reshape(-1))
            cols.append(Cc.reshape(-1))
            vals.append((-dt * theta * Mf).reshape(-1))
    M = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(lay.n, lay.n)).tocsr()
    M.sum_duplicates()
    M.sort_indices()
    return M


def rhs_matrix(coords, cells, cell_p1, nP2, nP1, params, dt, theta, K_cells=None):
    """Matrix of the up_-linear part of b, rhs of F at t_n
    (mpettotalpressuresolver.py:270-274, the rows w_i): for network
    rows f >= 1, Bm[f, g] M (the storage and exchange coefficients of
    pressure_coefficients) plus dt (1 - theta) K_f S on the diagonal block.
    The u and p_0 rows are zero.  Passing a subset of the cells gives the
    contributions of those cells only (rows whose cells all lie in the
    subset are then exact)."""
    A = int(params["J"])
    lay = Layout(nP2, nP1, A)
    _, lmbda = lame(params["E"], params["nu"])
    _, Bm = pressure_coefficients(params["alpha"], params["c"], params["S"], lmbda, dt, theta)
    et = element_tensors(coords, cells)
    nC = cells.shape[0]
    K = K_cells if K_cells is not None else np.tile(np.asarray(params["K"], float), (nC, 1))
    rows, cols, vals = [], [], []
    for f in range(1, A + 1):
        for g in range(A + 1):
            v = Bm[f, g] * et["M"]
            if f == g:
                v = v + dt * (1 - theta) * K[:, f - 1, None, None] * et["S"]
            R, Cc = _blk(lay.pdofs(cell_p1, f), lay.pdofs(cell_p1, g), v)
            rows.append(R.reshape(-1))
            cols.append(Cc.reshape(-1))
            vals.append(v.reshape(-1))
    return sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(lay.n, lay.n)).tocsr()


# ---------------------------------------------------------------------------
# two-field formulation (MPETSolver)
# ---------------------------------------------------------------------------


def two_field_pressure_coefficients(c, S, dt, theta):
    """A x A coefficients of the P1 mass in the pressure rows of the
    two-field form (mpet/mpetsolver.py:198, 201): Cm for lhs(F), Bm for the
    up_ part of rhs(F)."""
    A = len(c)
    Cm = np.zeros((A, A))
    Bm = np.zeros((A, A))
    for i in range(A):
        for j in range(A):
            if i == j:
                tr = sum(S[i][k] for k in range(A) if k != i)
                Cm[i, j] = -c[i] - dt * theta * tr
                Bm[i, j] = -c[i] + dt * (1 - theta) * tr
            else:
                Cm[i, j] = dt * theta * S[i][j]
                Bm[i, j] = -dt * (1 - theta) * S[i][j]
    return Cm, Bm


def assemble_two_field(coords, cells, cell_p2, cell_p1, nP2, nP1, params,
                       dt, theta, K_cells=None, robin=None, which="A"):
    """Global FP64 CSR of the two-field MPETSolver system, fields u,
    p_1..p_A with F = A pressure fields per node (mpet/mpetsolver.py:196-201,
    stress sigma = 2 mu eps + lambda div I of mpetproblem.py:18-24).

    which == "A": lhs(F) + sum_i lhs(F2a_i) (:252-253, added at :336-338, :413-415):
      (u,v)     int 2 mu eps(u):eps(v) + lambda div u div v       :196
      (p_i,v)   -alpha_i int p_i div v                            :197
      (u,q_i)   -alpha_i int div u q_i                            :199
      (p_j,q_i) -c_i delta_ij M - dt theta delta_ij K_i S
                - dt theta (delta_ij sum_k S_ik - S_ij) M          :198, 200-201
                - dt theta delta_ij int_{ds_i(2)} beta_i p_i q_i   :252
    which == "P": prec (:269-271): mu grad u : grad v;
      (c_i + dt theta sum_{j != i} S_ij) M + dt theta K_i S (the reference's
      ``dt*theta**inner(...)`` read as the product dt*theta*inner(...)).
    """
    A = int(params["J"])
    mu, lmbda = lame(params["E"], params["nu"])
    alpha, c, S = params["alpha"], params["c"], params["S"]
    lay = Layout(nP2, nP1, A, A)
    et = element_tensors(coords, cells)
    nC = cells.shape[0]
    if K_cells is None:
        K_cells = np.tile(np.asarray(params["K"], float), (nC, 1))
    ud = lay.udofs(cell_p2)
    blocks = []
    if which == "A":
        blocks.append((ud, ud, 2 * mu * et["E"] + lmbda * et["D"]))
        Cm, _ = two_field_pressure_coefficients(c, S, dt, theta)
        for i in range(A):
            pdi = lay.pdofs(cell_p1, i)
            blocks.append((ud, pdi, -alpha[i] * et["B"]))
            blocks.append((pdi, ud, -alpha[i] * np.transpose(et["B"], (0, 2, 1))))
            for j in range(A):
                v = Cm[i, j] * et["M"]
                if i == j:
                    v = v - dt * theta * K_cells[:, i, None, None] * et["S"]
                blocks.append((pdi, lay.pdofs(cell_p1, j), v))
    else:
        G = mu * et["G"]
        Gu = np.zeros((nC, 10, 3, 10, 3))
        for j in range(3):
            Gu[:, :, j, :, j] = G
        blocks.append((ud, ud, Gu.reshape(nC, 30, 30)))
        for i in range(A):
            a_i = c[i] + sum(dt * theta * S[i][j] for j in range(A) if j != i)
            v = a_i * et["M"] + dt * theta * K_cells[:, i, None, None] * et["S"]
            pdi = lay.pdofs(cell_p1, i)
            blocks.append((pdi, pdi, v))
    rows, cols, vals = [], [], []
    for (ld, rd, v) in blocks:
        R, Cc = _blk(ld, rd, v)
        rows.append(R.reshape(-1))
        cols.append(Cc.reshape(-1))
        vals.append(v.reshape(-1))
    if which == "A" and robin is not None:
        for i, rb in enumerate(robin):
            if rb is None:
                continue
            tri_dofs, Mf = rb
            pdi = lay.nu + A * tri_dofs + i
            R, Cc = _blk(pdi, pdi, Mf)
            rows.append(R.reshape(-1))
            cols.append(Cc.reshape(-1))
            vals.append((-dt * theta * Mf).reshape(-1))
    M = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(lay.n, lay.n)).tocsr()
    M.sum_duplicates()
    M.sort_indices()
    return M


def two_field_rhs_matrix(mesh, params, dt, theta, K_cells, lay):
    """Matrix of the up_-linear part of rhs(F) of the two-field form at t_n
    (:196-201): pressure rows i get -alpha_i int div u_ q_i,
    Bm_ij (M p_j_) and dt (1 - theta) K_i S p_i_."""
    m = mesh
    A = int(params["J"])
    alpha = params["alpha"]
    _, Bm = two_field_pressure_coefficients(params["c"], params["S"], dt, theta)
    et = element_tensors(m["coords"], m["cells"])
    nC = m["cells"].shape[0]
    K = K_cells if K_cells is not None else np.tile(np.asarray(params["K"], float), (nC, 1))
    ud = lay.udofs(m["cell_p2"])
    rows, cols, vals = [], [], []
    for i in range(A):
        pdi = lay.pdofs(m["cell_p1"], i)
        blocks = [(pdi, ud, -alpha[i] * np.transpose(et["B"], (0, 2, 1)))]
        for j in range(A):
            v = Bm[i, j] * et["M"]
            if i == j:
                v = v + dt * (1 - theta) * K[:, i, None, None] * et["S"]
            blocks.append((pdi, lay.pdofs(m["cell_p1"], j), v))
        for (ld, rd, v) in blocks:
            R, Cc = _blk(ld, rd, v)
            rows.append(R.reshape(-1))
            cols.append(Cc.reshape(-1))
            vals.append(v.reshape(-1))
    return sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(lay.n, lay.n)).tocsr()


# ---------------------------------------------------------------------------
# sources (Expression(degree=k) interpolated into P_k, then exact quadrature)
# ---------------------------------------------------------------------------


def cell_source_vector(coords, cells, cell_test, f, degree, t, ncomp):
    """int f . v over cells for P2 (cell_test (nC,10)) or P1 (nC,4) test
    functions, with f first interpolated into P_degree on each cell.
    Returns (nC, ntest, ncomp) local contributions."""
    J, det, _ = cell_geometry(coords, cells)
    nodes = lagrange_nodes(degree)
    v0 = coords[cells[:, 0]]
    X = v0[:, None, :] + np.einsum("cij,nj->cni", J, nodes)      # nC,nk,3
    fv = np.asarray(f(X.reshape(-1, 3), t), float).reshape(X.shape[0], X.shape[1], -1)
    order = degree + 2 + 1
    qp, qw = tet_rule(order)
    Lk = lagrange_eval(degree, qp) if degree > 0 else np.ones((len(qp), 1))
    T = p2_ref_values(qp) if cell_test.shape[1] == 10 else p1_ref_values(qp)
    fq = np.einsum("qm,cmk->cqk", Lk, fv)
    return np.einsum("q,c,qa,cqk->cak", qw, det, T, fq)[:, :, :ncomp]


def scatter_vector(n, dofs, vals):
    b = np.zeros(n)
    np.add.at(b, dofs.reshape(-1), vals.reshape(-1))
    return b


def facet_geometry(coords, tris):
    v = coords[tris]
    e1 = v[:, 1] - v[:, 0]
    e2 = v[:, 2] - v[:, 0]
    cr = np.cross(e1, e2)
    area2 = np.linalg.norm(cr, axis=1)          # = 2 * area
    return v, e1, e2, area2


def facet_p1_mass(coords, tris):
    """(nF,3,3) P1 mass on triangles: |F|/12 (1 + delta)."""
    _, _, _, a2 = facet_geometry(coords, tris)
    base = np.array([[2, 1, 1], [1, 2, 1], [1, 1, 2]], float) / 12.0
    return (a2 / 2.0)[:, None, None] * base


def facet_source_vector(coords, tris, tri_test_nodes, f, degree, t, normals=None):
    """int_F f . v over facets.  tri_test_nodes: 3 (P1) or 6 (P2, vertices then
    edges (1,2),(0,2),(0,1)) node coordinates are implied by tris.
    f(X, t) or f(X, t, normals) -> (n, ncomp)."""
    v, e1, e2, a2 = facet_geometry(coords, tris)
    kn = lagrange_nodes(degree, 2)
    X = v[:, 0][:, None, :] + kn[None, :, 0, None] * e1[:, None, :] + kn[None, :, 1, None] * e2[:, None, :]
    nF, nk = X.shape[:2]
    if normals is None:
        fv = np.asarray(f(X.reshape(-1, 3), t), float)
    else:
        fv = np.asarray(f(X.reshape(-1, 3), t, np.repeat(normals, nk, axis=0)), float)
    fv = fv.reshape(nF, nk, -1)
    qp, qw = tri_rule(degree + 2 + 1)
    Lk = lagrange_eval(degree, qp, 2) if degree > 0 else np.ones((len(qp), 1))
    x, y = qp[:, 0], qp[:, 1]
    lam = np.stack([1 - x - y, x, y], axis=1)
    if tri_test_nodes == 3:
        T = lam
    else:
        T = np.concatenate([lam * (2 * lam - 1),
                            np.stack([4 * lam[:, 1] * lam[:, 2], 4 * lam[:, 0] * lam[:, 2],
                                      4 * lam[:, 0] * lam[:, 1]], axis=1)], axis=1)
    fq = np.einsum("qm,fmk->fqk", Lk, fv)
    return np.einsum("q,f,qa,fqk->fak", qw, a2, T, fq)


# ---------------------------------------------------------------------------
# boundary conditions
# ---------------------------------------------------------------------------


def dirichlet_dofs(mesh_facets, facet_cell, facet_local, cells, cell_p2, cell_p1,
                   markers, field, lay):
    """DOFs in the closure of facets marked 0 (DOLFIN topological method).
    field 0 -> all u DOFs of the facet's 3 vertex + 3 edge P2 nodes;
    field k >= 1 -> the DOFs of layout pressure field k - 1 at its 3 vertices
    (sub-space k of the mixed space)."""
    sel = np.nonzero(np.asarray(markers) == 0)[0]
    if sel.size == 0:
        return np.zeros(0, dtype=np.int64)
    cl = facet_cell[sel]
    lf = facet_local[sel]
    verts = np.array(FACES)[lf]                                   # nF,3
    if field == 0:
        nodes = [cell_p2[cl, verts[:, k]] for k in range(3)]
        for (a, b) in ((1, 2), (0, 2), (0, 1)):
            pairs = np.sort(np.stack([verts[:, a], verts[:, b]], 1), axis=1)
            eidx = np.array([EDGES.index(tuple(p)) for p in pairs.tolist()])
            nodes.append(cell_p2[cl, 4 + eidx])
        nodes = np.unique(np.concatenate(nodes))
        return np.sort((3 * nodes[:, None] + np.arange(3)).reshape(-1))
    nodes = np.unique(np.concatenate([cell_p1[cl, verts[:, k]] for k in range(3)]))
    return np.sort(lay.nu + lay.F * nodes + (field - 1))


# ---------------------------------------------------------------------------
# nullspace multipliers (R-space Lagrange multipliers)
# ---------------------------------------------------------------------------


def rigid_motions(coords, cells):
    """mpet/rm_basis_L2.py:10-74, 3D branch: centroid c = int x / |Omega|,
    rotational Gram matrix R_ij = int (x-c) x e_i . (x-c) x e_j (:41-48),
    eigenpairs (w_k, v_k) of R (:51-53, numpy eigh, eigenvectors as rows);
    translations v_k / sqrt|Omega| (:58), rotations (x-c) x v_k / sqrt(w_k)
    (:64-72).  Integrals by degree-2 quadrature (exact for these integrands).
    Returns (list of six callables X -> (n, 3), volume)."""
    J, det, _ = cell_geometry(coords, cells)
    qp, qw = tet_rule(2)
    v0 = coords[cells[:, 0]]
    X = v0[:, None, :] + np.einsum("cij,qj->cqi", J, qp)           # nC,nq,3
    w = det[:, None] * qw[None, :]
    vol = w.sum()
    c = np.einsum("cq,cqi->i", w, X) / vol
    R = np.zeros((3, 3))
    E = np.eye(3)
    for i in range(3):
        for j in range(3):
            ci = np.cross(X - c, E[i])
            cj = np.cross(X - c, E[j])
            R[i, j] = np.einsum("cq,cqk,cqk->", w, ci, cj)
    eigw, eigv = np.linalg.eigh(R)
    eigv = eigv.T
    out = []
    for v in eigv:
        out.append(lambda P, v=v: np.broadcast_to(v / math.sqrt(vol), (len(P), 3)).copy())
    for v, ew in zip(eigv, eigw):
        out.append(lambda P, v=v, ew=ew: np.cross(np.asarray(P) - c, v) / math.sqrt(ew))
    return out, vol


def constraint_rows(coords, cells, cell_p2, cell_p1, lay, Z, fields):
    """Rows of the multiplier constraints as a sparse (m, n) matrix:
    int Z_k . v over the P2 vector test functions for each rigid motion Z_k
    (mpettotalpressuresolver.py:287-288, mpetsolver.py:205-206), then
    int q over the P1 test functions of each constrained pressure field
    (layout field number; :297-299, mpetsolver.py:212-214).  Element
    quadrature of degree 3 (exact: linear Z times quadratic phi)."""
    J, det, _ = cell_geometry(coords, cells)
    qp, qw = tet_rule(3)
    v0 = coords[cells[:, 0]]
    X = v0[:, None, :] + np.einsum("cij,qj->cqi", J, qp)
    phi = p2_ref_values(qp)
    psi = p1_ref_values(qp)
    nC = cells.shape[0]
    rows = []
    for Zk in Z:
        Zq = np.asarray(Zk(X.reshape(-1, 3)), float).reshape(nC, len(qp), 3)
        loc = np.einsum("q,c,qa,cqi->cai", qw, det, phi, Zq)          # nC,10,3
        dofs = 3 * cell_p2[:, :, None] + np.arange(3)[None, None, :]
        rows.append(scatter_vector(lay.n, dofs, loc))
    for f in fields:
        loc = np.einsum("q,c,qa->ca", qw, det, psi)
        rows.append(scatter_vector(lay.n, lay.pdofs(cell_p1, f), loc))
    return sp.csr_matrix(np.array(rows)) if rows else sp.csr_matrix((0, lay.n))


# ---------------------------------------------------------------------------
# a-posteriori error indicators (AdaptiveMPETSolver, two-field)
# ---------------------------------------------------------------------------


def p2_ref_hess():
    """(10, 3, 3) constant reference Hessians of the P2 basis."""
    dlam = np.array([[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    H = [4 * np.outer(dlam[i], dlam[i]) for i in range(4)]
    H += [4 * (np.outer(dlam[i], dlam[j]) + np.outer(dlam[j], dlam[i])) for (i, j) in EDGES]
    return np.array(H)


def error_indicators(m, params, x, xp, dt, f=None, fp=None, g=None, f_degree=0, g_degree=0,
                     t=0.0):
    """Cell indicators of mpet/adaptivempetsolver.py:119-153 for two-field
    states x (current) and xp (previous), layout F = A:
      zeta_p = sum_j h^2 ||RK_p,j||^2_K + sum_F avg(h)/2 ||RE_p,j||^2_F
      zeta_u = h^2 ||RK_u||^2_K + sum_F avg(h)/2 ||RE_u||^2_F,  zeta_u_dt alike
    with the residuals of :121-133 (h = longest edge, F interior facets) and
    the cell parts of R4p (:146-153).  Sources are callables (X, t) with
    Expression(degree) semantics (P_k interpolation per cell).  Cell
    integrals by Stroud conical-product quadrature, facet integrals by a
    collapsed Gauss rule; second derivatives from the reference Hessians.
    Returns (zeta_p, zeta_u, zeta_u_dt, r4) per cell."""
    A = int(params["J"])
    mu, lmbda = lame(params["E"], params["nu"])
    alpha = np.asarray(params["alpha"], float)
    c = np.asarray(params["c"], float)
    K = np.asarray(params["K"], float)
    S = np.asarray(params["S"], float)
    coords, cells, cp2 = m["coords"], m["cells"], m["cell_p2"]
    nC = cells.shape[0]
    nu = 3 * m["nP2"]
    lay = Layout(m["nP2"], m["nP1"], A, A)
    J, det, Jinv = cell_geometry(coords, cells)
    v = coords[cells]
    h = np.max([np.linalg.norm(v[:, i] - v[:, j], axis=1)
                for i in range(4) for j in range(i + 1, 4)], axis=0)
    vol = det / 6.0

    def ufield(z):
        return z[:nu].reshape(-1, 3)[cp2]                        # nC,10,3

    def pfield(z):
        return z[nu:lay.n].reshape(-1, A)[cells]                 # nC,4,A

    Href = p2_ref_hess()
    # physical Hessian: J^-T H_ref J^-1
    Hph = np.einsum("cki,akl,clj->caij", Jinv, Href, Jinv)      # nC,10,3,3
    g1 = np.einsum("ckm,ak->cam", Jinv, P1_REF_GRADS)           # nC,4,3

    def divsig(U):
        lap = np.einsum("caii,cak->ck", Hph, U)
        gd = np.einsum("caij,caj->ci", Hph, U)
        return mu * lap + (mu + lmbda) * gd

    u, up_ = ufield(x), ufield(xp)
    p, pp = pfield(x), pfield(xp)
    gp = np.einsum("caj,cam->cjm", p, g1)                        # nC,A,3
    gpp = np.einsum("caj,cam->cjm", pp, g1)
    Cu = divsig(u) - np.einsum("j,cjm->cm", alpha, gp)
    Cup = divsig(up_) - np.einsum("j,cjm->cm", alpha, gpp)
    deg = 2 * max(f_degree if (f is not None or fp is not None) else 0,
                  g_degree if g is not None else 0, 1)
    qp, qw = tet_rule(deg)
    X = coords[cells[:, 0]][:, None, :] + np.einsum("cij,qj->cqi", J, qp)

    def interp(fun, k, ncomp):
        nodes = lagrange_nodes(k)
        Xn = coords[cells[:, 0]][:, None, :] + np.einsum("cij,nj->cni", J, nodes)
        vals = np.asarray(fun(Xn.reshape(-1, 3), t), float).reshape(nC, len(nodes), -1)
        L = lagrange_eval(k, qp) if k > 0 else np.ones((len(qp), 1))
        return np.einsum("qm,cmk->cqk", L, vals)[:, :, :ncomp]

    fq = interp(f, f_degree, 3) if f is not None else np.zeros((nC, len(qp), 3))
    fpq = interp(fp, f_degree, 3) if fp is not None else np.zeros((nC, len(qp), 3))
    RK = fq + Cu[:, None, :]
    RKp = fpq + Cup[:, None, :]
    wq = det[:, None] * qw[None, :]
    zu = h ** 2 * np.einsum("cq,cqi,cqi->c", wq, RK, RK)
    Rd = (RK - RKp) / dt
    zudt = h ** 2 * np.einsum("cq,cqi,cqi->c", wq, Rd, Rd)
    # pressure residuals at the quadrature points
    psi = p1_ref_values(qp)                                      # nq,4
    pq = np.einsum("qa,caj->cqj", psi, p)
    ppq = np.einsum("qa,caj->cqj", psi, pp)
    gr = np.einsum("ckm,qak->cqam", Jinv, p2_ref_grads(qp))      # nC,nq,10,3
    divu = np.einsum("cqai,cai->cq", gr, u - up_)
    gq = interp(g, g_degree, A) if g is not None else np.zeros((nC, len(qp), A))
    RKp_ = (-gq - c[None, None, :] * (pq - ppq) / dt - alpha[None, None, :] * divu[:, :, None] / dt
            - np.einsum("ij,cqj->cqj", S, pq) + np.einsum("ij,cqi->cqj", S, pq))
    zp = h ** 2 * np.einsum("cq,cqj,cqj->c", wq, RKp_, RKp_)
    dpq = pq - ppq
    r4 = np.einsum("j,cjm,cjm->c", K, gp - gpp, gp - gpp) * vol
    for j in range(A):
        for i in range(A):
            if S[i, j] > 3.0e-16:
                r4 += np.einsum("cq,cq->c", wq, S[i, j] * (dpq[:, :, j] - dpq[:, :, i]) * dpq[:, :, j])
    # interior facets
    cf = m["cell_facets"].reshape(-1)
    ext = m["exterior"]
    slots = np.nonzero(~ext[cf])[0]
    order = np.argsort(cf[slots], kind="stable")
    pair = slots[order].reshape(-1, 2)
    tq, tw = tri_rule(2)
    for sa, sb in pair:
        ca, la, cb = sa // 4, sa % 4, sb // 4
        P = coords[cells[ca, list(FACES[la])]]
        e1, e2 = P[1] - P[0], P[2] - P[0]
        nrm = np.cross(e1, e2)
        area = 0.5 * np.linalg.norm(nrm)
        nrm /= np.linalg.norm(nrm)
        if np.dot(coords[cells[ca, la]] - P[0], nrm) > 0:
            nrm = -nrm
        havg = 0.5 * (h[ca] + h[cb])
        Xq = P[0] + np.outer(tq[:, 0], e1) + np.outer(tq[:, 1], e2)
        wts = tw * 2.0 * area

        def local(cell, z):
            ref = (Xq - coords[cells[cell, 0]]) @ Jinv[cell].T       # nq,3 reference coords
            G = np.einsum("km,qak->qam", Jinv[cell], p2_ref_grads(ref))
            U = ufield(z)[cell]
            Gu = np.einsum("ai,qaj->qij", U, G)
            sig = mu * (Gu + np.transpose(Gu, (0, 2, 1))) + \
                lmbda * np.einsum("qii->q", Gu)[:, None, None] * np.eye(3)
            pv = p1_ref_values(ref) @ pfield(z)[cell]                # nq,A
            gpl = np.einsum("aj,am->jm", pfield(z)[cell], g1[cell])  # A,3
            return sig, pv, gpl

        res = []
        for z in (x, xp):
            sa_, pa_, ga_ = local(ca, z)
            sb_, pb_, gb_ = local(cb, z)
            RE = (np.einsum("j,qj->q", alpha, pa_ - pb_)[:, None] * nrm
                  - np.einsum("qij,j->qi", sa_ - sb_, nrm))
            REp = -K * ((ga_ - gb_) @ nrm)
            res.append((RE, REp))
        (RE, REp), (REo, _) = res
        Fu = havg * np.einsum("q,qi,qi->", wts, RE, RE)
        Fd = havg * np.einsum("q,qi,qi->", wts, (RE - REo) / dt, (RE - REo) / dt)
        Fp = havg * area * np.sum(REp ** 2)
        for cell in (ca, cb):
            zu[cell] += 0.5 * Fu
            zudt[cell] += 0.5 * Fd
            zp[cell] += 0.5 * Fp
    return zp, zu, zudt, r4


def bordered(A, B, C=None):
    """[A B^T; B C] (C = 0 when None)."""
    return sp.bmat([[A, B.T], [B, C]], format="csr")


def apply_row_replace(A, dofs):
    """DirichletBC.apply(A): rows replaced by identity rows (not symmetric)."""
    A = A.tolil(copy=True)
    for d in dofs:
        A.rows[d] = [d]
        A.data[d] = [1.0]
    return A.tocsr()


def apply_symmetric(A, dofs, b=None):
    """MatZeroRowsColumns(dofs, diag=1, x=b, b=b) (bc_symmetric.py:11-18):
    b_other -= A[other, dofs] b[dofs]; rows and columns zeroed, diagonal 1.
    b[dofs] keeps its values (diag * x = b)."""
    A = A.tocsr()
    n = A.shape[0]
    mask = np.zeros(n, bool)
    mask[dofs] = True
    if b is not None:
        xd = np.zeros(n)
        xd[dofs] = b[dofs]
        corr = A @ xd
        b[~mask] -= corr[~mask]
    keep = sp.diags((~mask).astype(float))
    A2 = (keep @ A @ keep).tocsr()
    A2 = A2 + sp.diags(mask.astype(float))
    A2 = A2.tocsr()
    A2.sort_indices()
    return A2


# ---------------------------------------------------------------------------
# MINRES (Paige & Saunders 1975), preconditioned
# ---------------------------------------------------------------------------


def minres(Aop, b, Minv, x0=None, rtol=1e-10, maxiter=10000):
    """Textbook preconditioned MINRES.  Stops when the preconditioned
    residual norm estimate phibar <= rtol * ||r0||_{M^-1}.  Returns
    (x, iterations, phibar history)."""
    n = b.shape[0]
    x = np.zeros(n) if x0 is None else x0.copy()
    r1 = b - Aop(x)
    y = Minv(r1)
    beta1 = math.sqrt(max(float(r1 @ y), 0.0))
    if beta1 == 0.0:
        return x, 0, [0.0]
    oldb, beta, dbar, epsln, phibar = 0.0, beta1, 0.0, 0.0, beta1
    cs, sn = -1.0, 0.0
    w = np.zeros(n)
    w2 = np.zeros(n)
    r2 = r1.copy()
    hist = [beta1]
    it = 0
    while it < maxiter:
        it += 1
        s = 1.0 / beta
        v = s * y
        y = Aop(v)
        if it >= 2:
            y = y - (beta / oldb) * r1
        alfa = float(v @ y)
        y = y - (alfa / beta) * r2
        r1 = r2
        r2 = y
        y = Minv(r2)
        oldb = beta
        beta = math.sqrt(max(float(r2 @ y), 0.0))
        oldeps = epsln
        delta = cs * dbar + sn * alfa
        gbar = sn * dbar - cs * alfa
        epsln = sn * beta
        dbar = -cs * beta
        gamma = max(math.hypot(gbar, beta), 1e-300)
        cs = gbar / gamma
        sn = beta / gamma
        phi = cs * phibar
        phibar = sn * phibar
        w1 = w2
        w2 = w
        w = (v - oldeps * w1 - delta * w2) / gamma
        x = x + phi * w
        hist.append(phibar)
        if phibar <= rtol * beta1:
            break
    return x, it, hist


# ---------------------------------------------------------------------------
# a complete problem description + time loop
# ---------------------------------------------------------------------------


class OracleProblem:
    """Raw-array restatement of MPETProblem + MPETTotalPressureSolver.

    mesh: dict with coords, cells, cell_p2, cell_p1, nP2, nP1, x2 (P2 node
          coordinates), facets, facet_cell, facet_local, exterior.
    params: {J, E, nu, alpha, c, K, S}; solver: {dt, theta, T, t}.
    f(X,t)->(n,3), g[i](X,t)->(n,), u_bar(X,t)->(n,3), p_bar[i](X,t)->(n,)
    (numpy callables); degrees as in Expression(degree=k).
    markers: momentum (nfacets,), continuity list of (nfacets,).
    """

    def __init__(self, mesh, params, solver, f=None, g=None, u_bar=None,
                 p_bar=None, f_degree=3, g_degree=3, K_cells=None,
                 momentum_markers=None, continuity_markers=None,
                 s=None, s_degree=3, I=None, I_degree=3, beta=None,
                 p_robin=None, robin_degree=1, p_robin_degree=None,
                 formulation="total_pressure", permc_spec="COLAMD", refine_steps=2,
                 u_nullspace=False, p_nullspace=None):
        if formulation not in ("total_pressure", "two_field"):
            raise ValueError(formulation)
        self.two_field = formulation == "two_field"
        # column ordering of SuperLU; COLAMD factorises these row-replaced
        # systems fastest (8^3 all-Dirichlet MMS: 8.6 s vs 133 s with MMD on
        # A + A^T; benchmark workloads: 12 s vs 16 s total-pressure, 9 s vs
        # 51 s two-field)
        self.permc_spec = permc_spec
        self.refine_steps = int(refine_steps)
        self.m = mesh
        self.params = params
        self.A = int(params["J"])
        self.dt = float(solver["dt"])
        self.theta = float(solver["theta"])
        self.T = float(solver["T"])
        self.t0 = float(solver.get("t", 0.0))
        self.f, self.g, self.u_bar, self.p_bar = f, g, u_bar, p_bar
        self.f_degree, self.g_degree = f_degree, g_degree
        self.s, self.s_degree = s, s_degree
        self.I, self.I_degree = I, I_degree
        self.beta, self.p_robin, self.robin_degree = beta, p_robin, robin_degree
        self.p_robin_degree = robin_degree if p_robin_degree is None else p_robin_degree
        self.K_cells = K_cells
        self.lay = Layout(mesh["nP2"], mesh["nP1"], self.A,
                          self.A if self.two_field else self.A + 1)
        nF = mesh["facets"].shape[0]
        self.mm = momentum_markers if momentum_markers is not None else np.full(nF, 2**63 - 1)
        self.cm = continuity_markers or [np.full(nF, 2**63 - 1) for _ in range(self.A)]
        self._robin_mass = self._robin_terms()
        m = mesh
        asm = assemble_two_field if self.two_field else assemble_matrix
        self.A_mat = asm(m["coords"], m["cells"], m["cell_p2"], m["cell_p1"],
                         m["nP2"], m["nP1"], params, self.dt, self.theta,
                         K_cells, robin=self._robin_mass, which="A")
        self.P_mat = asm(m["coords"], m["cells"], m["cell_p2"], m["cell_p1"],
                         m["nP2"], m["nP1"], params, self.dt, self.theta,
                         K_cells, which="P")
        self.bc_dofs = self._bc_dofs()
        self._B_mat = None
        # nullspace multipliers (mpettotalpressuresolver.py:179-202, 285-306;
        # mpetsolver.py:107-128, 203-215): p_nullspace[k] constrains the k-th
        # pressure trial function of the reference's mixed space, i.e. layout
        # field k (p_0 first in the total-pressure formulation)
        self.u_nullspace = bool(u_nullspace)
        self.p_fields = [k for k, f in enumerate(p_nullspace or []) if f]
        self.nmult = 0
        if self.u_nullspace or self.p_fields:
            Z = []
            if self.u_nullspace:
                Z, _ = rigid_motions(m["coords"], m["cells"])
            self.Z = Z
            self.C_mat = constraint_rows(m["coords"], m["cells"], m["cell_p2"], m["cell_p1"],
                                         self.lay, Z, self.p_fields)
            self.nmult = self.C_mat.shape[0]
            self.A_mat = bordered(self.A_mat, self.C_mat)

    # -- boundary facets -----------------------------------------------------
    def _marked_facets(self, markers, value):
        m = self.m
        sel = np.nonzero((np.asarray(markers) == value) & m["exterior"])[0]
        cl = m["facet_cell"][sel]
        lf = m["facet_local"][sel]
        verts = np.array(FACES)[lf]
        return sel, cl, verts

    def _robin_terms(self):
        if self.beta is None:
            return None
        out = []
        m = self.m
        for i in range(self.A):
            sel, cl, verts = self._marked_facets(self.cm[i], 2)
            if sel.size == 0:
                out.append(None)
                continue
            tri_nodes = np.stack([m["cell_p1"][cl, verts[:, k]] for k in range(3)], 1)
            tri_v = np.stack([m["cells"][cl, verts[:, k]] for k in range(3)], 1)
            # int_F beta(t0) psi psi with beta interpolated in P_robin_degree
            vals = facet_source_matrix(m["coords"], tri_v, self.beta[i],
                                       self.robin_degree, self.t0)
            out.append((tri_nodes, vals))
        return out

    def _bc_dofs(self):
        m = self.m
        out = [dirichlet_dofs(m["facets"], m["facet_cell"], m["facet_local"], m["cells"],
                              m["cell_p2"], m["cell_p1"], self.mm, 0, self.lay)]
        for i in range(self.A):
            out.append(dirichlet_dofs(m["facets"], m["facet_cell"], m["facet_local"],
                                      m["cells"], m["cell_p2"], m["cell_p1"],
                                      self.cm[i], self.lay.net(i) + 1, self.lay))
        return out

    def bc_values(self, t):
        """Nodal Dirichlet values at time t for each BC DOF set."""
        m = self.m
        lay = self.lay
        vals = []
        d0 = self.bc_dofs[0]
        if d0.size:
            nodes = d0 // 3
            comp = d0 % 3
            ub = self.u_bar(m["x2"][nodes], t) if self.u_bar is not None else np.zeros((nodes.size, 3))
            vals.append(np.asarray(ub)[np.arange(nodes.size), comp])
        else:
            vals.append(np.zeros(0))
        for i in range(self.A):
            d = self.bc_dofs[i + 1]
            if d.size:
                nodes = (d - lay.nu) // lay.F
                pb = self.p_bar[i](m["coords"][nodes], t) if self.p_bar is not None else np.zeros(nodes.size)
                vals.append(np.asarray(pb, float).reshape(-1))
            else:
                vals.append(np.zeros(0))
        return vals

    # -- right-hand side ------------------------------------------------------
    def B_matrix(self):
        """Matrix of the up_-linear part of b (rhs of F, evaluated at t_n)."""
        if self._B_mat is not None:
            return self._B_mat
        if self.two_field:
            self._B_mat = two_field_rhs_matrix(self.m, self.params, self.dt, self.theta,
                                               self.K_cells, self.lay)
            return self._B_mat
        m = self.m
        # beta(t_theta) of a Robin term is re-evaluated per step in rhs()
        self._B_mat = rhs_matrix(m["coords"], m["cells"], m["cell_p1"], m["nP2"], m["nP1"],
                                 self.params, self.dt, self.theta, self.K_cells)
        return self._B_mat

    def rhs(self, up_prev, t_n):
        """b assembled as in solve_iterative :461-483 (before BCs)."""
        m = self.m
        lay = self.lay
        b = self.B_matrix() @ up_prev                          # L at t_n
        t_theta = t_n + self.theta * self.dt
        t_np1 = t_theta + (1.0 - self.theta) * self.dt
        # L1: dt g_i w_i dx + dt I_i w_i ds_i(1) at t_theta
        if self.g is not None:
            for i in range(self.A):
                if self.g[i] is None:
                    continue
                loc = cell_source_vector(m["coords"], m["cells"], m["cell_p1"],
                                         lambda X, t, gi=self.g[i]: np.asarray(gi(X, t)).reshape(-1, 1),
                                         self.g_degree, t_theta, 1)
                b += self.dt * scatter_vector(lay.n, lay.pdofs(m["cell_p1"], lay.net(i)), loc[..., 0])
        if self.I is not None:
            for i in range(self.A):
                if self.I[i] is None:
                    continue
                sel, cl, verts = self._marked_facets(self.cm[i], 1)
                if sel.size == 0:
                    continue
                tri_v = np.stack([m["cells"][cl, verts[:, k]] for k in range(3)], 1)
                tri_n = np.stack([m["cell_p1"][cl, verts[:, k]] for k in range(3)], 1)
                loc = facet_source_vector(m["coords"], tri_v, 3,
                                          lambda X, t, Ii=self.I[i]: np.asarray(Ii(X, t)).reshape(-1, 1),
                                          self.I_degree, t_theta)
                b += self.dt * scatter_vector(lay.n, lay.pdofs(tri_n, lay.net(i)), loc[..., 0])
        # rhs(L2_i): dt beta((1-theta) p_i_ - p_robin) w_i ds_i(2) at t_theta
        if self.beta is not None:
            for i in range(self.A):
                sel, cl, verts = self._marked_facets(self.cm[i], 2)
                if sel.size == 0:
                    continue
                tri_v = np.stack([m["cells"][cl, verts[:, k]] for k in range(3)], 1)
                tri_n = np.stack([m["cell_p1"][cl, verts[:, k]] for k in range(3)], 1)
                pprev = up_prev[lay.pdofs(tri_n, lay.net(i))]                # nF,3

                def fdat(X, t, bi=self.beta[i], pr=self.p_robin[i] if self.p_robin else None):
                    return np.asarray(bi(X, t)).reshape(-1, 1)
                Mf = facet_source_matrix(m["coords"], tri_v, self.beta[i], self.robin_degree, t_theta)
                loc = (1 - self.theta) * np.einsum("fab,fb->fa", Mf, pprev)
                if self.p_robin is not None and self.p_robin[i] is not None:
                    # int beta_h p_robin_h w: each coefficient interpolated in
                    # its own P_k (FFC coefficient semantics), product exact
                    pr = facet_product_vector(m["coords"], tri_v, self.beta[i],
                                              self.robin_degree, self.p_robin[i],
                                              self.p_robin_degree, t_theta)
                    loc = loc - pr
                b += self.dt * scatter_vector(lay.n, lay.pdofs(tri_n, lay.net(i)), loc)
        # L0: f . v dx + s . v ds_m(1) at t_{n+1}
        if self.f is not None:
            loc = cell_source_vector(m["coords"], m["cells"], m["cell_p2"], self.f,
                                     self.f_degree, t_np1, 3)
            b += scatter_vector(lay.n, lay.udofs(m["cell_p2"]), loc.reshape(-1, 30))
        if self.s is not None:
            sel, cl, verts = self._marked_facets(self.mm, 1)
            if sel.size:
                tri_v = np.stack([m["cells"][cl, verts[:, k]] for k in range(3)], 1)
                nodes = [m["cell_p2"][cl, verts[:, k]] for k in range(3)]
                for (a, bb) in ((1, 2), (0, 2), (0, 1)):
                    pairs = np.sort(np.stack([verts[:, a], verts[:, bb]], 1), axis=1)
                    eidx = np.array([EDGES.index(tuple(p)) for p in pairs.tolist()])
                    nodes.append(m["cell_p2"][cl, 4 + eidx])
                nodes = np.stack(nodes, 1)                                    # nF,6
                normals = outward_normals(m["coords"], m["cells"], cl, verts)
                loc = facet_source_vector(m["coords"], tri_v, 6, self.s, self.s_degree,
                                          t_np1, normals=normals)              # nF,6,3
                dofs = (3 * nodes[:, :, None] + np.arange(3)).reshape(-1, 18)
                b += scatter_vector(lay.n, dofs, loc.reshape(-1, 18))
        return b, t_theta, t_np1

    # -- time loops -----------------------------------------------------------
    def solve_direct(self, up0):
        """Generator mirroring solve_direct (:330-411): LU of row-replaced A
        (bordered by the multiplier rows when the problem has nullspaces;
        up0 and the yielded states then hold the multipliers after the
        fields)."""
        A = self.A_mat.copy()
        alld = np.concatenate(self.bc_dofs)
        A = apply_row_replace(A, alld)
        lu = spla.splu(A.tocsc(), permc_spec=self.permc_spec)
        up_prev = up0.copy()
        t = self.t0
        n = self.lay.n
        while t < self.T - 1e-9:
            b, t_theta, t_np1 = self.rhs(up_prev[:n], t)
            if self.nmult:
                b = np.concatenate([b, np.zeros(self.nmult)])
            for d, v in zip(self.bc_dofs, self.bc_values(t_np1)):
                b[d] = v
            up = lu.solve(b)
            # iterative refinement in FP64 (two steps by default): the
            # two-field systems with tiny K are ill-conditioned enough that the
            # unrefined LU solution's weakest pressure field moves by ~1e-8
            # with the column ordering; the timed CPU baseline (one LU solve
            # per step, as the reference) uses none
            for _ in range(self.refine_steps):
                up += lu.solve(b - A @ up)
            yield up, t_np1
            up_prev = up
            t = t_np1

    def eliminated_system(self, b, t_np1):
        """Symmetric elimination of all BCs in order (:487-489)."""
        if self.nmult:
            raise NotImplementedError("the oracle solves bordered systems directly only")
        A = self.A_mat.copy()
        for d, v in zip(self.bc_dofs, self.bc_values(t_np1)):
            b[d] = v
            A = apply_symmetric(A, d, b)
        return A, b

    def preconditioner(self):
        if self.nmult:
            raise NotImplementedError("the oracle solves bordered systems directly only")
        P = self.P_mat.copy()
        for d in self.bc_dofs:
            P = apply_symmetric(P, d)
        return P

    def solve_iterative(self, up0, rtol=1e-10, exact_blocks=True):
        """Generator mirroring solve_iterative (:414-505): MINRES on the
        symmetrically eliminated system, preconditioned by the eliminated P.
        The block-diagonal P is inverted exactly (sparse LU per block) —
        the reference applies BoomerAMG instead, which the GPU path
        replaces by its own AMG; only the converged solution is compared."""
        P = self.preconditioner()
        lu = spla.splu(P.tocsc())
        up_prev = up0.copy()
        t = self.t0
        self.niter = []
        while t < self.T - 1e-9:
            b, t_theta, t_np1 = self.rhs(up_prev, t)
            Ael, bel = self.eliminated_system(b, t_np1)
            x, it, _ = minres(lambda v: Ael @ v, bel, lu.solve, x0=up_prev, rtol=rtol)
            self.niter.append(it)
            yield x, t_np1
            up_prev = x
            t = t_np1


def facet_source_matrix(coords, tris, beta, degree, t):
    """(nF,3,3) int_F beta psi_a psi_b with beta interpolated into P_degree."""
    v, e1, e2, a2 = facet_geometry(coords, tris)
    kn = lagrange_nodes(degree, 2)
    X = v[:, 0][:, None, :] + kn[None, :, 0, None] * e1[:, None, :] + kn[None, :, 1, None] * e2[:, None, :]
    nF, nk = X.shape[:2]
    bv = np.asarray(beta(X.reshape(-1, 3), t), float).reshape(nF, nk)
    qp, qw = tri_rule(degree + 2 + 1)
    Lk = lagrange_eval(degree, qp, 2) if degree > 0 else np.ones((len(qp), 1))
    x, y = qp[:, 0], qp[:, 1]
    lam = np.stack([1 - x - y, x, y], axis=1)
    bq = np.einsum("qm,fm->fq", Lk, bv)
    return np.einsum("q,f,qa,qb,fq->fab", qw, a2, lam, lam, bq)


def facet_product_vector(coords, tris, f1, k1, f2, k2, t):
    """(nF,3) int_F f1_h f2_h psi_a, f1 and f2 interpolated into P_k1 / P_k2
    on each facet, the product integrated exactly."""
    v, e1, e2, a2 = facet_geometry(coords, tris)
    vals = []
    for f, k in ((f1, k1), (f2, k2)):
        kn = lagrange_nodes(k, 2)
        X = v[:, 0][:, None, :] + kn[None, :, 0, None] * e1[:, None, :] \
            + kn[None, :, 1, None] * e2[:, None, :]
        nF, nk = X.shape[:2]
        vals.append(np.asarray(f(X.reshape(-1, 3), t), float).reshape(nF, nk))
    qp, qw = tri_rule(k1 + k2 + 2)
    L1 = lagrange_eval(k1, qp, 2) if k1 > 0 else np.ones((len(qp), 1))
    L2 = lagrange_eval(k2, qp, 2) if k2 > 0 else np.ones((len(qp), 1))
    x, y = qp[:, 0], qp[:, 1]
    lam = np.stack([1 - x - y, x, y], axis=1)
    q1 = np.einsum("qm,fm->fq", L1, vals[0])
    q2 = np.einsum("qm,fm->fq", L2, vals[1])
    return np.einsum("q,f,qa,fq,fq->fa", qw, a2, lam, q1, q2)


def outward_normals(coords, cells, cl, verts):
    v = coords[cells[cl]]                                   # nF,4,3
    tri = np.stack([v[np.arange(len(cl)), verts[:, k]] for k in range(3)], 1)
    n = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    cen = v.mean(axis=1)
    sgn = np.sign(np.einsum("fi,fi->f", n, tri[:, 0] - cen))
    return n * sgn[:, None]
