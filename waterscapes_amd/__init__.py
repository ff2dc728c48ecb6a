"""waterscapes_amd — MI355X-native backend of the total-pressure MPET time
step of meg-simula/waterscapes (`mpet`), and of the two-field MPETSolver.

``from waterscapes_amd import *`` exposes the same names the reference's
``from mpet import *`` does for this path (mpet/__init__.py:10-17), plus the
DOLFIN look-alikes (mesh, markers, Constant, Expression, Function) the API
is written against.
"""
from .mesh import (Mesh, BoxMesh, UnitCubeMesh, MeshFunction, FacetFunction, SubDomain,
                   CompiledSubDomain, DOLFIN_EPS, INVALID)
from .fem import (Constant, Expression, Function, FacetNormal, MixedSpace, interpolate, assign,
                  errornorm)
from .problem import MPETProblem, convert_to_mu_lmbda, convert_to_E_nu
from .solver import MPETTotalPressureSolver, DirichletBC, Parameters
from .mpetsolver import MPETSolver
from .krylov import SolverConfig
from .amg import AmgParams
from .io import File, XDMFFile, load_mesh, save_mesh

__all__ = ["Mesh", "BoxMesh", "UnitCubeMesh", "MeshFunction", "FacetFunction", "SubDomain",
           "CompiledSubDomain", "DOLFIN_EPS", "INVALID", "Constant", "Expression", "Function",
           "FacetNormal", "MixedSpace", "interpolate", "assign", "errornorm", "MPETProblem",
           "convert_to_mu_lmbda", "convert_to_E_nu", "MPETTotalPressureSolver", "MPETSolver",
           "DirichletBC",
           "Parameters", "SolverConfig", "AmgParams", "File", "XDMFFile", "load_mesh",
           "save_mesh"]
