"""MPETSolver: the two-field MPET formulation (fields u, p_1..p_A, no total
pressure) of mpet/mpetsolver.py on one MI355X.

Same constructor, parameters (default_params, :86-98), attributes (a,
a_robin, L, L0, L1, prec, up_, up, dt, bcs, solver_monitor, params) and
entry points: solve() (which, like the reference, always takes the direct
path, :300-314), solve_direct() (generator yielding (up, t) per step and
updating up_ after the yield, :382-462), solve_iterative() (:465-570) and
step(dt, up_) (one step from the current time, updating up and time but
not up_, :317-379).  The forms of :196-254 are fixed descriptors: the
matrices are expanded on the GPU from the assembled primitives
(`mpet_tf_expand_csr`).

The reference factorises A with MUMPS (row-replaced Dirichlet rows); there
is no sparse LU on the GPU, so both solve paths run FP64 MINRES on the
symmetrically eliminated system (scalar CSR operator, block AMG
preconditioner of prec) to the relative tolerance "krylov_rtol" (1e-12 by
default), which reproduces the LU solution to round-off level (tested on
the oracle).  MINRES needs a symmetric system: S must be symmetric.
"""
from __future__ import annotations

import time as _time

import numpy as np
import torch

from .amg import AmgParams
from .fem import Constant, Function, MixedSpace
from .krylov import SolverConfig
from .nullspace import pressure_fields
from .solver import DirichletBC, FormDescriptor, Parameters, StepEngine


class MPETSolver(object):
    """Two-field MPET solver (mpet/mpetsolver.py:21-570), see module docstring."""

    def __init__(self, problem, params=None):
        self.problem = problem
        self.params = self.default_params()
        if params is not None:
            self.params.update(params)
        self.solver_monitor = {}
        self._engine = None
        self.create_variational_forms()
        self.create_dirichlet_bcs()

    @staticmethod
    def default_params():
        """mpet/mpetsolver.py:86-98 plus backend keys: "krylov_rtol" (relative
        preconditioned-residual tolerance of the MINRES solve standing in for
        LU), "maxit", "u_smoother_degree", "u_level_degrees",
        "p_smoother_degree", "precision" ("fp64", or "mixed": FP32 MINRES
        and preconditioner on an FP32 copy of the CSR system inside FP64
        refinement to "outer_tol", measured like krylov_rtol; "inner_rtol"
        of the inner solves), "extrapolate" / "extrapolate_order" (initial
        guess of a step extrapolated in time from the previous states; the
        tolerance is relative to the right-hand side, so only the iteration
        count depends on it), "device"."""
        params = Parameters("MPETSolver")
        params.add("dt", 0.1)
        params.add("t", 0.0)
        params.add("T", 1.0)
        params.add("theta", 1.0)
        params.add("u_degree", 2)
        params.add("p_degree", 1)
        params.add("direct_solver", True)
        params.add("testing", False)
        params.add("krylov_rtol", 1e-12)
        params.add("maxit", 20000)
        params.add("u_smoother_degree", 3)
        params.add("u_level_degrees", (1,))
        params.add("p_smoother_degree", 2)
        params.add("precision", "fp64")
        params.add("outer_tol", 1e-12)
        params.add("inner_rtol", 1e-6)
        params.add("extrapolate", True)
        params.add("extrapolate_order", 4)
        params.add("device", "cuda")
        return params

    # -- set-up ---------------------------------------------------------------
    def create_function_spaces(self, mesh):
        """[P2]^3 x P1^A (:100-131), plus the R-space multipliers of the
        nullspace options (:107-128): a 6-vector for the rigid motions, one
        scalar per pressure with p_has_nullspace, after the field DOFs."""
        if self.params["u_degree"] != 2 or self.params["p_degree"] != 1:
            raise NotImplementedError("the GPU backend implements P2 displacement / P1 pressures")
        return MixedSpace(mesh, int(self.problem.params["J"]), total_pressure=False,
                          u_nullspace=bool(self.problem.u_has_nullspace),
                          p_fields=pressure_fields(self.problem))

    def create_variational_forms(self, include_preconditioner=False):
        """Fields and form descriptors of :134-298; returns (forms, prec,
        fields, dt) like the reference."""
        VQ = self.create_function_spaces(self.problem.mesh)
        A = VQ.A
        dev = torch.device(self.params["device"])
        self.up_ = Function(VQ, torch.zeros(VQ.dim(), dtype=torch.float64, device=dev))
        self.up = Function(VQ, torch.zeros(VQ.dim(), dtype=torch.float64, device=dev))
        self.dt = Constant(self.params["dt"])
        self.a = FormDescriptor("a", terms=":196-201 lhs")
        self.a_robin = [FormDescriptor(f"a_robin[{i}]", terms=":252-253") for i in range(A)]
        self.L = FormDescriptor("L", terms=":196-201 rhs")
        self.L0 = FormDescriptor("L0", terms=":233")
        self.L1 = [FormDescriptor(f"L1[{i}]", terms=":249, :254") for i in range(A)]
        self.prec = FormDescriptor("prec", terms=":269-271") if include_preconditioner else 0
        forms_ = (self.a, self.a_robin, self.L, self.L0, self.L1)
        return forms_, self.prec, (self.up_, self.up), self.dt

    def create_dirichlet_bcs(self):
        """[bcs0, bcs1] of :65-83: u on sub(0), p_i on sub(i+1)."""
        VP = self.up.function_space()
        bcs0 = [DirichletBC(VP, 0, self.problem.u_bar, self.problem.momentum_boundary_markers, 0)]
        bcs1 = []
        for i in range(int(self.problem.params["J"])):
            bcs1 += [DirichletBC(VP, i + 1, self.problem.p_bar[i],
                                 self.problem.continuity_boundary_markers[i], 0)]
        self.bcs = [bcs0, bcs1]
        return [bcs0, bcs1]

    def _config(self):
        p = self.params
        if p["precision"] not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if p["precision"] == "mixed" and self.up.function_space().nmult:
            raise ValueError("systems with nullspace multipliers are solved in FP64 only")
        return SolverConfig(mode=p["precision"], rtol=float(p["krylov_rtol"]),
                            outer_tol=float(p["outer_tol"]), inner_rtol=float(p["inner_rtol"]),
                            maxit=int(p["maxit"]),
                            amg=AmgParams(smoother_degree=int(p["u_smoother_degree"]),
                                          level_degrees=tuple(p["u_level_degrees"])),
                            pamg=AmgParams(smoother_degree=int(p["p_smoother_degree"])))

    def engine(self):
        """Device state for the current dt (A depends on dt; rebuilt when a
        step() call changes it)."""
        dt = float(self.params["dt"])
        if self._engine is None or self._engine.dt != dt:
            self._engine = StepEngine(self.problem, self.up_.function_space(), dt,
                                      self.params["theta"], self._config(),
                                      self.params["device"],
                                      extrapolate=bool(self.params["extrapolate"]),
                                      extrapolate_order=int(self.params["extrapolate_order"]))
        return self._engine

    # -- time stepping ----------------------------------------------------------
    def _record(self, eng, t0):
        torch.cuda.synchronize()
        st = eng.stats[-1]
        self.solver_monitor.setdefault("niter", []).append(int(st.inner_iterations))
        self.solver_monitor.setdefault("step_seconds", []).append(_time.perf_counter() - t0)

    def step(self, dt=None, up_=None):
        """One step from the current time t to t + dt (:317-379): updates up
        and the problem's time, not up_ (the reference reads up_ from the
        solver as well; the up_ argument is accepted and unused, as there)."""
        if dt is not None:
            self.params["dt"] = float(dt)
            self.dt.assign(dt)
        eng = self.engine()
        time = self.problem.time
        t0 = _time.perf_counter()
        x = self.up.data
        x.copy_(self.up_.data)
        eng.step(self.up_.data, x, float(time))
        self._record(eng, t0)

    def solve(self):
        """Like the reference (:300-314): always the direct path."""
        return self.solve_direct()

    def solve_direct(self):
        """Reference :382-462 (see module docstring for the GPU realisation)."""
        return self._run(testing=False)

    def solve_iterative(self):
        """Reference :465-570: MINRES with the block preconditioner prec;
        "testing" starts each solve from random values (:514-517)."""
        return self._run(testing=bool(self.params["testing"]))

    def _run(self, testing):
        T = self.params["T"]
        time = self.problem.time
        eng = self.engine()
        self.solver_monitor["niter"] = []
        self.solver_monitor["step_seconds"] = []
        x = self.up.data
        if testing:
            x.copy_(torch.randn(x.shape[0], dtype=torch.float64, device=x.device))
        else:
            self.up.assign(self.up_)
        while float(time) < (T - 1.0e-9):
            t0 = _time.perf_counter()
            eng.step(self.up_.data, x, float(time), testing_guess=testing)
            testing = False
            self._record(eng, t0)
            yield self.up, float(time)
            self.up_.assign(self.up)
        self.solver_monitor["A"] = eng.system
        self.solver_monitor["P"] = eng.krylov
