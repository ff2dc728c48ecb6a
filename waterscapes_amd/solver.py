"""MPETTotalPressureSolver: drop-in for mpet/mpettotalpressuresolver.py on one
MI355X.

Same constructor, parameters (default_params, :146-162), attributes (F, L0,
L1, L2, prec, up_, up, solver_monitor, params) and generator protocol
(solve() yields (up, t) per step and updates up_ after the yield,
:497-502).  The forms of :164-328 are not built symbolically: F, L0, L1, L2
and prec are descriptors of the fixed forms of the total-pressure
formulation, whose element tensors the HIP assembly kernel evaluates.

Per step, on the GPU (StepEngine.step):
  b  = rhs(F)(up_)             at t_n        (:461)   one P1-pattern kernel
  b += dt g_i, dt I_i, Robin   at t_theta    (:470-476)
  b += f, s                    at t_{n+1}    (:482-483)
  x0 = up_ with the Dirichlet values at t_{n+1} in place   (:487-488)
  x  = MINRES(A_elim, P_elim)(b_elim)        (:489-494)
The symmetric elimination of :487-489 is folded into the operators (rows
masked, lifting = the initial residual), so A is never copied (:457).

direct_solver=True (the reference default, LU with MUMPS, :330-411) has no
sparse LU on the GPU: it runs the same FP64 Krylov solve to rtol 1e-12,
which reproduces the LU solution to round-off level; row replacement and
symmetric elimination give the same solution (tested on the oracle).

Nullspaces (problem.u_has_nullspace, p_has_nullspace; :179-202, 285-306):
the multipliers of the reference's R spaces follow the field DOFs of up
(up.split() returns u, p_0..p_A, then the multipliers), and the bordered
system runs in FP64 MINRES (waterscapes_amd/nullspace.py).
"""
from __future__ import annotations

import math
import time as _time

import numpy as np
import torch

from . import forms, refel
from .device import DeviceSystem, TwoFieldSystem
from .facets import FacetSet, FacetTerm
from .fem import Constant, Function, MixedSpace
from .amg import AmgParams
from .krylov import KrylovSolver, SolverConfig
from .nullspace import Constraints, has_nullspace, pressure_fields
from .problem import resolved_params

NEUMANN_MARKER = 1
ROBIN_MARKER = 2


class Parameters(dict):
    """dolfin.Parameters look-alike: a dict with add()/update()."""

    def __init__(self, name="", **kw):
        super().__init__(**kw)
        self.name = name

    def add(self, key, value):
        self[key] = value


class FormDescriptor:
    """Stand-in for a UFL form of the total-pressure formulation."""

    def __init__(self, name, **info):
        self.name = name
        self.info = info

    def __repr__(self):
        return f"<form {self.name}>"


class DirichletBC:
    """DirichletBC(V.sub(i), value, markers, 0): the DOFs in the closure of
    facets marked 0 and their nodal values at the time of apply."""

    def __init__(self, space, field, value, markers, marker_value=0):
        if marker_value != 0:
            raise ValueError("only marker 0 denotes Dirichlet conditions")
        self.space = space
        self.field = field
        self.value = value
        nodes = forms.dirichlet_nodes(space, markers, field)
        self.nodes = nodes
        self.dofs = forms.field_dofs(space, field, nodes)
        dm = space.dofmap
        self.coords = dm.x2[nodes] if field == 0 else dm.x1[nodes]

    def values(self, t):
        if self.dofs.size == 0:
            return np.zeros(0)
        v = self.value.eval(self.coords, t)
        if self.field == 0:
            return np.broadcast_to(v, (self.nodes.size, 3)).reshape(-1)
        return np.broadcast_to(v, (self.nodes.size,)).astype(np.float64)

    def values_device(self, t, device):
        """values() as a device tensor, the nodal evaluation done where the
        node coordinates are kept (Expression.eval_device)."""
        if self.dofs.size == 0:
            return torch.zeros(0, dtype=torch.float64, device=device)
        if not hasattr(self.value, "eval_device"):
            return torch.as_tensor(np.array(self.values(t), dtype=np.float64)).to(device)
        if getattr(self, "_coords_dev", None) is None:
            self._coords_dev = torch.as_tensor(np.ascontiguousarray(self.coords)).to(device)
        v = self.value.eval_device(self._coords_dev, t).to(torch.float64)
        n = self.nodes.size
        if self.field == 0:
            return v.expand(n, 3).reshape(-1) if v.dim() > 1 else \
                v[:, None].expand(n, 3).reshape(-1)
        return v.expand(n).contiguous()

    def get_boundary_values(self, t=None):
        return dict(zip(self.dofs.tolist(), self.values(t).tolist()))


class _Source:
    """A volume source interpolated into P_k on each cell (Expression
    semantics) and integrated against P2 (u rows) or P1 (network rows)."""

    def __init__(self, engine, expr, test_degree, ncomp, scale, dof_base, dof_stride, dof_off):
        self.e = engine
        self.expr = expr
        self.k = int(getattr(expr, "degree", 0))
        mesh = engine.space.mesh
        x = mesh.coordinates()
        c = mesh.cells()
        bary = refel.pk_nodes_bary(self.k)                           # nk,4
        self.X = np.einsum("nk,ckd->cnd", bary, x[c])                # nC,nk,3
        # the P_k nodes of every cell stay in HBM: time-dependent sources are
        # evaluated there each step (Expression.eval_device), not on the host
        self.Xd = torch.as_tensor(self.X.reshape(-1, 3)).to(engine.device)
        self.Q = torch.as_tensor(refel.source_weights(test_degree, self.k)).to(engine.device)
        self.test_degree = test_degree
        self.ncomp = ncomp
        self.scale = scale
        self.dof = (dof_base, dof_stride, dof_off)
        self._cache = None

    def add(self, b, t):
        if self._cache is None or self.expr.depends_on_time():
            npts = self.Xd.shape[0]
            if hasattr(self.expr, "eval_device"):
                v = self.expr.eval_device(self.Xd, t)
            else:                      # Constant and other host-only coefficients
                v = torch.as_tensor(np.asarray(self.expr.eval(self.X.reshape(-1, 3), t),
                                               dtype=np.float64)).to(self.e.device)
            shape = (npts,) + ((self.ncomp,) if self.ncomp > 1 else ())
            self._cache = v.to(torch.float64).expand(shape).contiguous()
        self.e.system.cell_source(self.test_degree, self.Q, self._cache, self.ncomp, self.scale,
                                  *self.dof, b)


def extrapolated_guess(hist, order, up_prev, x):
    """x = the polynomial extrapolation in time of degree k = min(len(hist),
    order) through u_n = up_prev and the previous states hist (newest
    first): sum_j (-1)^j C(k+1, j+1) u_{n-j}; x = u_n when k = 0.  Exact
    for states that are polynomials of degree k in time."""
    k = min(len(hist), order)
    x.copy_(up_prev)
    if k:
        x.mul_(float(k + 1))
        for j in range(1, k + 1):
            x.add_(hist[j - 1], alpha=float((-1) ** j * math.comb(k + 1, j + 1)))


def adaptive_order(hist, order, up_prev, nu, F, nP1, tmp):
    """The degree k in 1 .. min(order, len(hist) - 1) whose extrapolation
    would have predicted the newest state up_prev best from the states
    before it (max over the fields - u and the F pressures - of the relative
    L2 error); min(order, len(hist)) when the history is too short to
    compare.  tmp: a scratch vector like up_prev."""
    kmax = min(order, len(hist) - 1)
    if kmax < 1:
        return min(order, len(hist))
    scale = _field_norms(up_prev, nu, F, nP1)
    best, best_err = kmax, None
    for k in range(1, kmax + 1):
        extrapolated_guess(hist[1:], k, hist[0], tmp)
        tmp.sub_(up_prev)
        err = (_field_norms(tmp, nu, F, nP1) / scale.clamp_min(1e-300)).max()
        if best_err is None or bool(err < best_err):
            best, best_err = k, err
    return best


def _field_norms(x, nu, F, nP1):
    """L2 norms of the u part and of each of the F node-interleaved pressure
    fields of x (entries past them, multipliers, are left out)."""
    pu = torch.linalg.vector_norm(x[:nu]).reshape(1)
    pp = torch.linalg.vector_norm(x[nu:nu + F * nP1].view(nP1, F), dim=0)
    return torch.cat([pu, pp])


def push_history(hist, order, up_prev):
    """Keep up_prev as the newest of at most `order` previous states."""
    if order < 1:
        return
    if len(hist) < order:
        hist.append(torch.empty_like(up_prev))
    hist.insert(0, hist.pop())          # the oldest buffer receives u_n
    hist[0].copy_(up_prev)


def _is_zero(e):
    return bool(getattr(e, "is_zero", lambda: False)())


def robin_facets(problem, space, device):
    """Robin facets of every network (markers 2 of the continuity boundary
    markers, :324) with a nonzero beta: [(i, FacetSet)] and the matching
    left-hand-side entries [(i, FacetSet, beta, None)]."""
    robin, lhs = [], []
    for i in range(space.A):
        if _is_zero(problem.beta[i]):
            continue
        fs = FacetSet(space, problem.continuity_boundary_markers[i], ROBIN_MARKER,
                      torch.device(device))
        if fs.n:
            robin.append((i, fs))
            lhs.append((i, fs, problem.beta[i], None))
    return robin, lhs


def batch_system(problem, space, dt, theta, device="cuda", assemble=True):
    """The DeviceSystem a batched StepEngine of this instance uses (no
    structured copies, no CSR); assemble=False defers the assembly to a
    batched one (device.assemble_batch)."""
    if not space.total_pressure:
        raise NotImplementedError("batches of two-field instances")
    params, cellK = resolved_params(problem)
    return DeviceSystem(space, params, dt, theta, K_cells=cellK, device=device,
                        build_csr=False, structured=False,
                        robin=robin_facets(problem, space, device)[1], assemble=assemble)


class StepEngine:
    """Device-resident state of one MPET instance and its time step
    (system: an already built DeviceSystem of a batched engine, from
    batch_system)."""

    def __init__(self, problem, space, dt, theta, config: SolverConfig, device="cuda",
                 extrapolate=False, batched=False, extrapolate_order=4, system=None,
                 extrapolate_adaptive=False):
        self.problem = problem
        self.batched = bool(batched)
        self.extrapolate = bool(extrapolate)
        self.ex_order = int(extrapolate_order)
        # adaptive: each step extrapolates with the degree (up to ex_order)
        # that best predicted the step just taken (one state more kept)
        self.ex_adaptive = bool(extrapolate_adaptive)
        self.ex_used = []
        if not 1 <= self.ex_order <= 6:
            raise ValueError("extrapolate_order must be 1 .. 6")
        # order 4 (quartic through five states): the 48^3 benchmark's
        # initial residual drops from 1e-4 (linear) to 4e-8 of ||b||, and
        # the steps from ~227 to ~112 MINRES iterations (cubic: ~153); the
        # outer tolerance fixes the error whatever the guess.  Orders 5 and
        # 6 cut the 80-step run's mean further (144 -> 131 -> 123
        # iterations) but cost more in the first steps after a start-up
        # transient (their coefficients, summing to 2^(k+1) - 1, amplify
        # whatever the early states do not share with a polynomial); 4 stays
        # the default
        self.hist = []                  # previous states u_{n-1}, u_{n-2}, .. (newest first)
        self.hist_t = []                # their times
        self.space = space
        self.device = torch.device(device)
        self.dt, self.theta = float(dt), float(theta)
        params, cellK = resolved_params(problem)
        self.params = params
        A = space.A
        self.nullspace = has_nullspace(problem)
        if self.nullspace and batched:
            raise NotImplementedError("batches of instances with nullspace multipliers")
        # exterior-facet terms (markers 1 = Neumann, 2 = Robin, :309-324)
        self.t0 = float(problem.time)
        self.robin, robin_lhs = robin_facets(problem, space, self.device)
        # batched: the batch packs the operator and hierarchies of all its
        # instances, so no per-instance structured copies or solver handle
        _t0 = _time.perf_counter()
        if system is not None:
            if not batched:
                raise ValueError("a prebuilt system is for batched engines")
            self.system = system
        elif space.total_pressure:
            self.system = DeviceSystem(space, params, dt, theta, K_cells=cellK, device=device,
                                       build_csr=(not batched and config.mode == "fp64"
                                                  and config.use_csr is not False),
                                       structured=not batched, robin=robin_lhs)
        else:
            if batched:
                raise NotImplementedError("batches of two-field instances")
            self.system = TwoFieldSystem(space, params, dt, theta, K_cells=cellK,
                                         device=device, robin=robin_lhs)
        # the two-field solver (mpet/mpetsolver.py:362-374, :428-448) sets
        # t_{n+1} = t_n + dt directly; the total-pressure one t_theta + (1-theta) dt
        self.t1_from_t0 = not space.total_pressure
        # Dirichlet conditions (create_dirichlet_bcs, :123-144; two-field
        # mpetsolver.py:65-83, networks on sub(i+1))
        self.bcs = [DirichletBC(space, 0, problem.u_bar, problem.momentum_boundary_markers, 0)]
        for i in range(A):
            self.bcs.append(DirichletBC(space, space.network_sub(i), problem.p_bar[i],
                                        problem.continuity_boundary_markers[i], 0))
        self.bc_dofs_host = np.concatenate([bc.dofs for bc in self.bcs]).astype(np.int64)
        self.bc_dofs = torch.as_tensor(self.bc_dofs_host).to(self.device)
        self.constraints = None
        if self.nullspace:
            if space.nmult != 6 * bool(problem.u_has_nullspace) + len(pressure_fields(problem)):
                raise ValueError("the function space's multipliers do not match the "
                                 "problem's nullspace flags")
            self.constraints = Constraints(self.system, bool(problem.u_has_nullspace),
                                           pressure_fields(problem))
        _t1 = _time.perf_counter()
        self.krylov = KrylovSolver(self.system, [bc.dofs for bc in self.bcs], config,
                                   create_handle=not batched, constraints=self.constraints)
        self.config = config
        self.setup_seconds = {"system": _t1 - _t0, "krylov": _time.perf_counter() - _t1}
        # sources
        nu, F = space.nu, space.F
        self.f_src = None
        if not problem.f.is_zero():
            self.f_src = _Source(self, problem.f, 2, 3, 1.0, 0, 3, 0)
        self.g_src = []
        for i in range(A):
            gi = problem.g[i]
            if not gi.is_zero():
                self.g_src.append(_Source(self, gi, 1, 1, self.dt, nu, F,
                                          space.network_sub(i) - 1))
        # facet sources: traction s on ds_m(1) (P2 test, 3 components), fluxes
        # I_i on ds_i(1) and Robin data on ds_i(2) (P1 test, network i rows)
        self.s_term = None
        if not _is_zero(problem.s):
            fs = FacetSet(space, problem.momentum_boundary_markers, NEUMANN_MARKER, self.device)
            if fs.n:
                self.s_term = FacetTerm(fs, 2, 0, 3, 0, 3)
        self.I_terms = []
        for i in range(A):
            if _is_zero(problem.I[i]):
                continue
            fs = FacetSet(space, problem.continuity_boundary_markers[i], NEUMANN_MARKER,
                          self.device)
            if fs.n:
                self.I_terms.append((i, FacetTerm(fs, 1, nu, F, space.network_sub(i) - 1, 1)))
        self.robin_terms = []
        for i, fs in self.robin:
            fi = space.network_sub(i) - 1
            gdof = torch.as_tensor((nu + F * fs.p1 + fi).astype(np.int32)).to(self.device)
            self.robin_terms.append((i, FacetTerm(fs, 1, nu, F, fi, 1), gdof))
        # multiplier rows of b (after the ndofs field rows) stay 0: the
        # constraints are homogeneous (the forms have no multiplier rhs)
        self.b = torch.zeros(space.dim(), dtype=torch.float64, device=self.device)
        self.stats = []

    def bc_values(self, t=None):
        """Dirichlet values of all BC DOFs (device tensor, in bc_dofs order)."""
        if not self.bcs:
            return torch.zeros(0, dtype=torch.float64, device=self.device)
        return torch.cat([bc.values_device(t, self.device) for bc in self.bcs])

    def assemble_rhs(self, up_prev, t_n):
        """b at the reference's time levels (:461-483), moving the problem's
        time Constant through t_theta to t_{n+1} exactly as :467-479 do, so
        Expressions bound to it see the same values.  Returns t_{n+1}."""
        b = self.b
        time = self.problem.time
        prob = self.problem
        dt, theta = float(self.dt), self.theta
        self.system.rhs_pressure(up_prev, b)
        t_theta = float(t_n) + theta * dt
        time.assign(t_theta)
        for src in self.g_src:                      # dt g_i w_i dx          (:322)
            src.add(b, None)
        for i, term in self.I_terms:                # dt I_i w_i ds_i(1)     (:322)
            term.add_source(b, prob.I[i], dt)
        for i, term, gdof in self.robin_terms:      # rhs(L2_i) on ds_i(2)   (:324)
            if theta != 1.0:
                term.add_robin_prev(b, prob.beta[i], up_prev, gdof, dt * (1.0 - theta))
            if not _is_zero(prob.p_robin[i]):
                term.add_product(b, prob.p_robin[i], prob.beta[i], -dt)
        t_np1 = float(t_n) + dt if self.t1_from_t0 else float(time) + (1.0 - theta) * dt
        time.assign(t_np1)
        if self.f_src is not None:                  # f . v dx               (:313)
            self.f_src.add(b, None)
        if self.s_term is not None:                 # s . v ds_m(1)          (:313)
            self.s_term.add_source(b, prob.s, 1.0)
        return t_np1

    def reset_history(self):
        self.hist = []
        self.hist_t = []

    def _valid_history(self, t_n):
        """Drop the stored states that are not u_{n-1}, u_{n-2}, .. at the
        uniform spacing dt behind t_n (a step repeated from the same state,
        a changed dt, a jump in time): the extrapolation weights assume
        that spacing."""
        dt = self.dt
        k = 0
        while k < len(self.hist_t) and abs(self.hist_t[k] - (t_n - (k + 1) * dt)) <= \
                1e-9 * max(dt, 1e-300):
            k += 1
        del self.hist[k:]
        del self.hist_t[k:]

    def step(self, up_prev, x, t_n, testing_guess=False):
        """One implicit step from up_prev (device FP64) into x; returns t_{n+1}."""
        t_np1 = self.assemble_rhs(up_prev, t_n)
        g = self.bc_values(None)
        if self.extrapolate:
            self._valid_history(float(t_n))
        if not testing_guess:
            order = self.ex_order
            if self.extrapolate and self.ex_adaptive:
                order = adaptive_order(self.hist, self.ex_order, up_prev, self.space.nu,
                                       self.space.A + 1, self.space.dofmap.nP1, x)
                self.ex_used.append(order)
            extrapolated_guess(self.hist if self.extrapolate else [], order, up_prev, x)
        if self.extrapolate:
            keep = self.ex_order + (1 if self.ex_adaptive else 0)
            push_history(self.hist, keep, up_prev)
            self.hist_t.insert(0, float(t_n))
            del self.hist_t[keep:]
        if self.bc_dofs.numel():
            x.index_copy_(0, self.bc_dofs, g)
        st = self.krylov.solve(self.b, x)
        self.stats.append(st)
        return t_np1, st


class MPETTotalPressureSolver(object):
    """Total-pressure MPET solver (mpet/mpettotalpressuresolver.py:15-512).

    Backend parameters added to default_params: "precision" ("fp64",
    "mixed" = FP32 MINRES/AMG in FP64 refinement, or "fp64-pc32" = FP64
    MINRES with the FP32 preconditioner), "krylov_rtol", "outer_tol",
    "inner_rtol", "maxit", "p0_weight" ("robust" / "reference"),
    "extrapolate" (initial guess extrapolated in time from up_ and the
    previous states instead of up_, polynomial degree "extrapolate_order",
    default 4; None = on for "mixed" and the direct path,
    on for "mixed", whose outer tolerance bounds the error independently of
    the guess, off otherwise), "u_smoother_degree" / "p_smoother_degree"
    (Chebyshev degrees of the two AMG hierarchies), "u_level_degrees" /
    "p_level_degrees" (their degrees on levels 0, 1, ..; 0 or absent: the
    smoother degree),
    "u_max_coarse" (rows below which the displacement hierarchy ends in a
    dense solve), "device"."""

    def __init__(self, problem, params=None):
        self.problem = problem
        self.monitor = False
        self.solver_monitor = {}
        self.params = self.default_params()
        if params is not None:
            self.params.update(params)
        F, L0, L1, L2, prec, up_, up = self.create_variational_forms()
        self.F = F
        self.L0 = L0
        self.L1 = L1
        self.L2 = L2
        self.prec = prec
        self.up_ = up_
        self.up = up
        self._engine = None
        self._engine_key_built = None

    @staticmethod
    def default_params():
        """mpet/mpettotalpressuresolver.py:146-162 plus backend keys."""
        params = Parameters("MPETSolver")
        params.add("dt", 0.05)
        params.add("t", 0.0)
        params.add("T", 1.0)
        params.add("theta", 0.5)
        params.add("u_degree", 2)
        params.add("p_degree", 1)
        params.add("direct_solver", True)
        params.add("testing", False)
        params.add("precision", "fp64")
        params.add("krylov_rtol", 1e-10)
        params.add("outer_tol", 3e-11)
        params.add("inner_rtol", 1e-6)
        params.add("maxit", 5000)
        params.add("p0_weight", "robust")
        params.add("extrapolate", None)
        params.add("extrapolate_order", 4)
        # pick the degree (up to extrapolate_order) per step: the one that
        # best predicted the previous step
        params.add("extrapolate_adaptive", False)
        params.add("u_smoother_degree", 3)
        params.add("p_smoother_degree", 2)
        params.add("u_level_degrees", (1,))
        params.add("p_level_degrees", ())
        params.add("u_max_coarse", 2500)
        # plain FP64 path: "csr" = the scalar CSR matrix of A (a PETSc AIJ
        # analogue, default), "structured" = the node-pattern operator the
        # mixed path uses in FP64; host checks of the MINRES state every
        # check_every iterations (0: 1 for FP64, 10 for mixed) and chunks
        # replayed from a captured hipGraph (None: mixed only)
        params.add("fp64_operator", "csr")
        # operator of the inner FP32 MINRES of the mixed path: "cellwise" =
        # integrated cell by cell from the geometry (matrix-free u block),
        # "structured" = the FP32 copy of the assembled node-pattern operator
        params.add("fp32_operator", "cellwise")
        params.add("check_every", 0)
        params.add("use_graphs", None)
        # launch shapes from the autotuner (None: for systems of at least
        # 200,000 unknowns)
        params.add("autotune", None)
        params.add("device", "cuda")
        return params

    def create_dirichlet_bcs(self):
        """[bcs0, bcs1] as in :123-144 (DirichletBC look-alikes)."""
        VP = self.up.function_space()
        bcs0 = [DirichletBC(VP, 0, self.problem.u_bar, self.problem.momentum_boundary_markers, 0)]
        bcs1 = []
        for i in range(int(self.problem.params["J"])):
            bcs1 += [DirichletBC(VP, i + 2, self.problem.p_bar[i],
                                 self.problem.continuity_boundary_markers[i], 0)]
        return [bcs0, bcs1]

    def create_variational_forms(self):
        if self.params["u_degree"] != 2 or self.params["p_degree"] != 1:
            raise NotImplementedError("the GPU backend implements P2 displacement / P1 pressures")
        mesh = self.problem.mesh
        A = int(self.problem.params["J"])
        VW = MixedSpace(mesh, A, u_nullspace=bool(self.problem.u_has_nullspace),
                        p_fields=pressure_fields(self.problem))
        dev = torch.device(self.params["device"])
        up_ = Function(VW, torch.zeros(VW.dim(), dtype=torch.float64, device=dev))
        up = Function(VW, torch.zeros(VW.dim(), dtype=torch.float64, device=dev))
        F = FormDescriptor("F", terms=":267-274")
        L0 = FormDescriptor("L0", terms=":313")
        L1 = [FormDescriptor(f"L1[{i}]", terms=":322") for i in range(A)]
        L2 = [FormDescriptor(f"L2[{i}]", terms=":324") for i in range(A)]
        prec = FormDescriptor("prec", terms=":279-283") if not self.params["direct_solver"] else 0
        return F, L0, L1, L2, prec, up_, up

    def _config(self, direct):
        p = self.params
        if direct or (has_nullspace(self.problem) and p["precision"] != "fp64"):
            # the bordered systems of nullspace multipliers run in FP64 only
            return SolverConfig(mode="fp64", rtol=1e-12, maxit=int(p["maxit"]),
                                p0_weight=p["p0_weight"])
        if p["fp64_operator"] not in ("csr", "structured"):
            raise ValueError("fp64_operator must be 'csr' or 'structured'")
        if p["fp32_operator"] not in ("cellwise", "structured"):
            raise ValueError("fp32_operator must be 'cellwise' or 'structured'")
        return SolverConfig(mode=p["precision"], rtol=float(p["krylov_rtol"]),
                            use_csr=(p["fp64_operator"] == "csr") if p["precision"] == "fp64"
                            else None,
                            check_every=int(p["check_every"]), use_graphs=p["use_graphs"],
                            autotune=p["autotune"],
                            cell_operator=(p["fp32_operator"] == "cellwise"),
                            outer_tol=float(p["outer_tol"]), inner_rtol=float(p["inner_rtol"]),
                            maxit=int(p["maxit"]), p0_weight=p["p0_weight"],
                            amg=AmgParams(smoother_degree=int(p["u_smoother_degree"]),
                                          level_degrees=tuple(p["u_level_degrees"]),
                                          max_coarse=int(p["u_max_coarse"])),
                            pamg=AmgParams(smoother_degree=int(p["p_smoother_degree"]),
                                           level_degrees=tuple(p["p_level_degrees"])))

    def _engine_key(self, direct):
        """Everything the device engine is built from that a caller may
        change between runs: the solve path and precision, tolerances,
        smoother settings, dt, theta and the initial-guess policy."""
        p = self.params
        keys = ("precision", "krylov_rtol", "outer_tol", "inner_rtol", "maxit", "p0_weight",
                "extrapolate", "extrapolate_order", "extrapolate_adaptive", "u_smoother_degree",
                "p_smoother_degree",
                "u_level_degrees", "p_level_degrees", "u_max_coarse", "fp64_operator",
                "fp32_operator",
                "check_every", "use_graphs", "autotune", "dt", "theta", "device")
        return (bool(direct),) + tuple(repr(p[k]) for k in keys)

    def engine(self, direct=None):
        """The device engine of the requested solve path.  It is rebuilt when
        the path (direct / iterative) or a parameter it depends on changed
        since the last call, so solve_direct() after solve_iterative() runs
        the direct path's FP64 solve, not the cached iterative one."""
        direct = self.params["direct_solver"] if direct is None else direct
        key = self._engine_key(direct)
        if self._engine is None or self._engine_key_built != key:
            self._engine = None
            cfg = self._config(direct)
            ex = self.params["extrapolate"]
            # default: on for the mixed path and for the direct path (whose
            # tolerance is relative to the right-hand side, so the guess only
            # moves the iteration count); the plain FP64 iterative path keeps
            # the reference's guess up_ (:452-454)
            ex = (cfg.mode == "mixed" or direct) if ex is None else bool(ex)
            self._engine = StepEngine(self.problem, self.up_.function_space(),
                                      self.params["dt"], self.params["theta"], cfg,
                                      self.params["device"], extrapolate=ex,
                                      extrapolate_order=int(self.params["extrapolate_order"]),
                                      extrapolate_adaptive=bool(self.params["extrapolate_adaptive"]))
            self._engine_key_built = key
        return self._engine

    def _run(self, direct):
        dt = self.params["dt"]
        T = self.params["T"]
        time = self.problem.time
        eng = self.engine(direct)
        eng.reset_history()
        self.solver_monitor["niter"] = []
        self.solver_monitor["outer"] = []
        self.solver_monitor["residual"] = []
        self.solver_monitor["step_seconds"] = []
        x = self.up.data
        testing = bool(self.params["testing"]) and not direct
        if testing:
            x.copy_(torch.randn(x.shape[0], dtype=torch.float64, device=x.device))
        else:
            self.up.assign(self.up_)
        while float(time) < (T - 1.0e-9):
            t0 = _time.perf_counter()
            t = eng.step(self.up_.data, x, float(time), testing_guess=testing)[0]
            testing = False
            torch.cuda.synchronize()
            st = eng.stats[-1]
            self.solver_monitor["niter"] += [int(st.inner_iterations)]
            self.solver_monitor["outer"] += [int(st.outer_iterations)]
            self.solver_monitor["residual"] += [float(st.residual_norm)]
            self.solver_monitor["step_seconds"] += [_time.perf_counter() - t0]
            yield self.up, float(time)
            self.up_.assign(self.up)
            time.assign(t)
        self.solver_monitor["P"] = eng.krylov
        self.solver_monitor["A"] = eng.system

    def solve_direct(self):
        """Reference :330-411 (see module docstring for the GPU realisation)."""
        return self._run(True)

    def solve_iterative(self):
        """Reference :414-505."""
        return self._run(False)

    def solve(self):
        if self.params["direct_solver"]:
            return self.solve_direct()
        else:
            return self.solve_iterative()
