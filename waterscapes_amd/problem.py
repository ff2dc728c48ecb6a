"""MPETProblem: problem description of the MPET equations.

Mirror of mpet/mpetproblem.py:26-180 (same attribute names, defaults and
marker conventions): Dirichlet = 0, Neumann = 1, Robin = 2, unmarked facets
INVALID = sys.maxsize.
"""
from __future__ import annotations

import sys

import numpy as np

from .fem import Constant
from .mesh import MeshFunction


def convert_to_E_nu(mu, lmbda):
    """mpet/mpetproblem.py:8-11."""
    E = mu * (3 * lmbda + 2 * mu) / (lmbda + mu)
    nu = lmbda / (2 * (lmbda + mu))
    return (E, nu)


def convert_to_mu_lmbda(E, nu):
    """mpet/mpetproblem.py:13-16."""
    mu = E / (2.0 * ((1.0 + nu)))
    lmbda = nu * E / ((1.0 - 2.0 * nu) * (1.0 + nu))
    return (mu, lmbda)


class MPETProblem(object):
    """Find u and p_1..p_J such that

      - div(sigma(u) - sum_j alpha_j p_j I) = f
      c_j (p_j)_t + alpha_j div(u_t) - div K_j grad p_j + sum_i S_ji (p_j - p_i) = g_j

    with sigma(u) = 2 mu eps(u) + lambda div(u) I (mpet/mpetproblem.py:26-101).

    Attributes set by the user (defaults as in the reference, :141-149):
    f, s (vector), g, I, beta, p_robin (lists of scalars), u_bar, p_bar;
    momentum_boundary_markers and continuity_boundary_markers[j] facet
    functions (default INVALID); u_has_nullspace, p_has_nullspace.
    ``params['K'][j]`` may be a number / Constant or a per-cell array (DG0
    permeability field) of length mesh.num_cells().
    """

    def __init__(self, mesh, time, params=None):
        self.mesh = mesh
        self.time = time
        self.params = self.default_parameters()
        if params is not None:
            self.params.update(params)
        Js = range(int(self.params["J"]))
        gdim = self.mesh.geometry().dim()
        self.f = Constant((0.0,) * gdim)
        self.s = Constant((0.0,) * gdim)
        self.g = [Constant(0.0) for i in Js]
        self.I = [Constant(0.0) for i in Js]
        self.beta = [Constant(0.0) for i in Js]
        self.p_robin = [Constant(0.0) for i in Js]
        self.u_bar = Constant((0.0,) * gdim)
        self.p_bar = [Constant(0.0) for i in Js]
        INVALID = sys.maxsize
        tdim = mesh.topology().dim()
        markers = MeshFunction("size_t", mesh, tdim - 1)
        markers.set_all(INVALID)
        self.momentum_boundary_markers = markers
        self.continuity_boundary_markers = []
        for i in Js:
            markers = MeshFunction("size_t", mesh, tdim - 1)
            markers.set_all(INVALID)
            self.continuity_boundary_markers += [markers]
        self.u_has_nullspace = False
        self.p_has_nullspace = list(False for i in Js)

    @classmethod
    def default_parameters(cls):
        """mpet/mpetproblem.py:168-180 (J defaults to 1.0; S must be given)."""
        ps = {"J": 1.0,
              "rho": Constant(1.0),
              "nu": Constant(0.479),
              "E": Constant(1500),
              "alpha": (Constant(1.0),),
              "K": (Constant(1.0),),
              "X": ((Constant(1.0),),),
              "c": (Constant(1.0),)}
        return ps


def scalar(v):
    """float() of a number or Constant."""
    return float(v)


def resolved_params(problem):
    """Numeric parameter dict {J, E, nu, alpha, c, S, K-or-cells} of a problem;
    per-cell K arrays are returned separately as (nC, J)."""
    p = problem.params
    J = int(p["J"])
    if "S" not in p:
        raise KeyError("S")          # the reference reads params['S'] (:248)
    out = dict(J=J, E=scalar(p["E"]), nu=scalar(p["nu"]),
               alpha=tuple(scalar(a) for a in p["alpha"])[:J],
               c=tuple(scalar(a) for a in p["c"])[:J],
               S=tuple(tuple(scalar(x) for x in row)[:J] for row in p["S"])[:J])
    nC = problem.mesh.num_cells()
    K = p["K"]
    cellK = np.zeros((nC, J))
    uniform = []
    for j in range(J):
        kj = K[j]
        if isinstance(kj, MeshFunction):
            kj = kj.array()
        if np.ndim(kj) == 0 or isinstance(kj, Constant):
            cellK[:, j] = scalar(kj)
            uniform.append(scalar(kj))
        else:
            arr = np.asarray(kj, dtype=np.float64).reshape(-1)
            if arr.size != nC:
                raise ValueError(f"K[{j}] must be a scalar or have one value per cell")
            cellK[:, j] = arr
            uniform.append(None)
    out["K"] = tuple(u if u is not None else float("nan") for u in uniform)
    return out, cellK
