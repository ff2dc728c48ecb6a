"""Nullspace multipliers of MPET problems without enough Dirichlet conditions.

With ``problem.u_has_nullspace`` the reference constrains the displacement
to be L2-orthogonal to the rigid motions, and with ``p_has_nullspace[k]`` the
pressure field p[k] to have zero mean, through R-space Lagrange multipliers
(mpet/mpettotalpressuresolver.py:179-202, 285-306; mpet/mpetsolver.py:
107-128, 203-215).  The system becomes the bordered

    [A  B^T] [x]   [b]
    [B   0 ] [z] = [0]

with one dense row of B per multiplier.  Here the rigid motions are built on
the host from the mesh geometry (a 3x3 eigenproblem, as
mpet/rm_basis_L2.py:10-74), the rows of B on the device
(``mpet_constraint_rows``, from the P2 and P1 mass matrices), and the
bordered system is solved by the device MINRES of
``mpet_solver_create_bordered``.

The reference indexes the constrained pressure as p[k] over the pressure
trial functions of its mixed space: p_0, p_1, .. in the total-pressure
solver (so p_has_nullspace[0] constrains the total pressure p_0) and
p_1, .. in the two-field solver.  ``pressure_fields`` keeps that meaning.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr


class RigidMotions:
    """The six L2-orthonormal rigid motions of mpet/rm_basis_L2.py:10-74 (3D):
    translations v_k / sqrt(|Omega|) along the eigenvectors v_k of the
    rotational Gram matrix R_ij = int (x - c) x e_i . (x - c) x e_j, then the
    rotations (x - c) x v_k / sqrt(w_k) about the same axes through the
    centroid c (w_k the eigenvalues).  Integrals are exact (vertex formulas
    for polynomials of degree <= 2 on tetrahedra)."""

    def __init__(self, coords, cells):
        v = np.asarray(coords, dtype=np.float64)[np.asarray(cells)]     # nC,4,3
        J = v[:, 1:] - v[:, :1]
        vol_c = np.abs(np.linalg.det(J)) / 6.0
        self.volume = float(vol_c.sum())
        self.centroid = (vol_c[:, None] * v.mean(axis=1)).sum(axis=0) / self.volume
        y = v - self.centroid
        ys = y.sum(axis=1)
        # int y y^T over a tetrahedron = |K| / 20 (sum_k y_k y_k^T + ys ys^T)
        S = np.einsum("c,cki,ckj->ij", vol_c / 20.0, y, y) + \
            np.einsum("c,ci,cj->ij", vol_c / 20.0, ys, ys)
        R = np.trace(S) * np.eye(3) - S
        w, V = np.linalg.eigh(R)
        self.eigenvalues = w
        self.axes = V.T.copy()                  # rows: eigenvectors (eigv.T)

    def values(self, X):
        """(6, npts, 3) values of the six rigid motions at points X (npts, 3)."""
        X = np.asarray(X, dtype=np.float64)
        out = np.empty((6, X.shape[0], 3))
        s = 1.0 / np.sqrt(self.volume)
        for k in range(3):
            out[k] = self.axes[k] * s
        d = X - self.centroid
        for k in range(3):
            out[3 + k] = np.cross(d, self.axes[k]) / np.sqrt(self.eigenvalues[k])
        return out


def pressure_fields(problem):
    """Layout field numbers of the pressures with a constant nullspace."""
    return [k for k, flag in enumerate(problem.p_has_nullspace) if flag]


def has_nullspace(problem):
    return bool(problem.u_has_nullspace) or any(problem.p_has_nullspace)


class Constraints:
    """Device rows B of the multipliers of one instance (fields of the space
    in layout order) and the P2 mass for the preconditioner's u block."""

    def __init__(self, system, u_nullspace, fields):
        if not torch.cuda.is_available():
            raise RuntimeError("Constraints need a GPU (no CPU fallback)")
        L = _lib.lib()
        s = system
        dev = s.device
        self.fields = [int(f) for f in fields]
        self.nz = 6 if u_nullspace else 0
        self.m = self.nz + len(self.fields)
        if self.m < 1:
            raise ValueError("no nullspace to constrain")
        mesh = s.space.mesh
        self.rm = RigidMotions(mesh.coordinates(), s.space.dofmap.cell_p1)
        self.omega = self.rm.volume
        self.M2 = None
        Zt = None
        if u_nullspace:
            self.M2 = torch.empty(s.nnz["22"], dtype=torch.float64, device=dev)
            check(L.mpet_assemble_p2_mass(C.byref(s.mesh_desc), C.byref(s.pat), ptr(self.M2),
                                          _lib.stream_ptr()))
            Z = self.rm.values(s.space.dofmap.x2)                      # 6,nP2,3
            Zt = torch.as_tensor(np.ascontiguousarray(Z.reshape(6, -1))).to(dev)
        self.B = torch.empty(self.m * s.n, dtype=torch.float64, device=dev)
        fl = np.ascontiguousarray(self.fields, dtype=np.int32)
        check(L.mpet_constraint_rows(C.byref(s.pat), ptr(self.M2) if self.M2 is not None else None,
                                     ptr(s.prim_t["M1"]), s.nP2, s.nP1, s.F,
                                     ptr(Zt) if Zt is not None else None, self.nz,
                                     fl.ctypes.data if fl.size else None, len(self.fields),
                                     ptr(self.B), _lib.stream_ptr()))
        torch.cuda.current_stream().synchronize()      # Zt is freed on return
        self.n = s.n

    def rows(self):
        """B as an (m, n) view."""
        return self.B.view(self.m, self.n)
