"""Edge cases of the total-pressure path on the GPU: the smallest and
largest network counts (A = 1 and A = 7, i.e. F = 2 and F = 8 pressure
fields, the most the kernels are instantiated for), the one-cube mesh,
a zero right-hand side, and the rejections of out-of-range inputs."""
import numpy as np
import pytest

from conftest import cfg2_params, lognormal_K, oracle_mesh
from oracle import mpet_oracle as O
from test_gpu_assembly import _build, _csr_host, _entrywise
from test_gpu_solver import _oracle_step, _setup

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("A", [1, 7])
def test_network_count_extremes_assembly(A):
    params = cfg2_params(A)
    n = 2
    K = lognormal_K(6 * n ** 3, params["K"], seed=4)
    space, dev = _build(n, params, K_cells=K, dt=0.0125, theta=1.0)
    m = oracle_mesh(space)
    A_ref = O.assemble_matrix(m["coords"], m["cells"], m["cell_p2"], m["cell_p1"], m["nP2"],
                              m["nP1"], params, 0.0125, 1.0, K_cells=K, which="A")
    _entrywise(_csr_host(dev.csr), A_ref)


@pytest.mark.parametrize("A,mode,kw", [(1, "mixed", dict(outer_tol=1e-11)),
                                       (7, "mixed", dict(outer_tol=1e-11)),
                                       (7, "fp64", dict(rtol=1e-12))])
def test_network_count_extremes_solve(A, mode, kw):
    import torch
    params, space, dev, solver, dofs, K = _setup(n=3, A=A, mode=mode, seed=A, **kw)
    rng = np.random.default_rng(A)
    up_prev = rng.standard_normal(dev.n)
    g = np.zeros(dev.n)
    alld = np.concatenate(dofs)
    g[alld] = rng.standard_normal(alld.size)
    ref, _ = _oracle_step(space, params, K, dofs, up_prev, g)
    up_t = torch.tensor(up_prev, device="cuda")
    b = torch.empty_like(up_t)
    dev.rhs_pressure(up_t, b)
    x = up_t.clone()
    x[torch.as_tensor(alld, device="cuda")] = torch.tensor(g[alld], device="cuda")
    st = solver.solve(b, x)
    got = x.cpu().numpy()
    for f in range(space.A + 2):
        idx = space.field_index(f)
        e = np.linalg.norm(got[idx] - ref[idx]) / max(np.linalg.norm(ref[idx]), 1e-300)
        assert e <= 1e-8, (A, mode, f, e, st.inner_iterations)


def test_one_cube_mesh_and_zero_rhs():
    """n = 1 (6 tetrahedra, most DOFs on Dirichlet faces) solves to the
    oracle; a zero right-hand side with zero Dirichlet data returns exactly
    zero after no iterations."""
    import torch
    params, space, dev, solver, dofs, K = _setup(n=1, A=2, mode="fp64", rtol=1e-12)
    rng = np.random.default_rng(9)
    up_prev = rng.standard_normal(dev.n)
    g = np.zeros(dev.n)
    alld = np.concatenate(dofs)
    g[alld] = rng.standard_normal(alld.size)
    ref, _ = _oracle_step(space, params, K, dofs, up_prev, g)
    up_t = torch.tensor(up_prev, device="cuda")
    b = torch.empty_like(up_t)
    dev.rhs_pressure(up_t, b)
    x = up_t.clone()
    x[torch.as_tensor(alld, device="cuda")] = torch.tensor(g[alld], device="cuda")
    solver.solve(b, x)
    got = x.cpu().numpy()
    assert np.linalg.norm(got - ref) <= 1e-8 * np.linalg.norm(ref)
    z = torch.zeros(dev.n, dtype=torch.float64, device="cuda")
    xz = torch.zeros_like(z)
    st = solver.solve(z, xz)
    assert st.inner_iterations == 0 and st.converged
    assert torch.count_nonzero(xz).item() == 0


def test_out_of_range_inputs_are_rejected():
    import ctypes as C
    import torch
    from waterscapes_amd import _lib
    from waterscapes_amd.fem import MixedSpace
    from waterscapes_amd.mesh import UnitCubeMesh
    from waterscapes_amd.device import DeviceSystem
    # eight networks exceed the nine fields the kernels are built for
    with pytest.raises(Exception):
        DeviceSystem(MixedSpace(UnitCubeMesh(1, 1, 1), 8), cfg2_params(7) | dict(
            J=8, alpha=(0.25,) * 8, c=(1e-5,) * 8, K=(1e-6,) * 8,
            S=[[0.0] * 8 for _ in range(8)]), 0.01, 1.0, build_csr=True)
    # an operator descriptor without nonzeros is refused before any launch
    params, space, dev, solver, dofs, K = _setup(n=1, A=2, mode="fp64")
    d = dev.op_desc(0)
    d.pat.nnz21 = 0
    x = torch.zeros(dev.n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    rc = _lib.lib().mpet_op_apply(C.byref(d), _lib.ptr(x), None, _lib.ptr(y), _lib.stream_ptr())
    assert rc == 1


@pytest.mark.parametrize("precision,direct,extra", [("fp64", True, {}),
                                                    ("mixed", False, dict(outer_tol=1e-11))])
def test_unstructured_numbering_matches_oracle(precision, direct, extra):
    """A cube mesh with randomly permuted vertex and cell numbering and
    randomly moved interior vertices (generic Mesh: no box fast path,
    scattered DOF maps and colouring): the cfg1 manufactured-solution steps
    match the oracle's LU to 1e-8 per field."""
    from conftest import CFG1
    from test_gpu_api import build_problem, oracle_steps, field_errors
    from waterscapes_amd.mesh import Mesh, UnitCubeMesh
    base = UnitCubeMesh(3, 3, 3)
    rng = np.random.default_rng(3)
    X = base.coordinates().copy()
    inner = np.all((X > 1e-12) & (X < 1.0 - 1e-12), axis=1)
    X[inner] += (0.2 / 3) * rng.uniform(-1.0, 1.0, (int(inner.sum()), 3))
    perm = rng.permutation(X.shape[0])
    inv = np.argsort(perm)
    cells = inv[base.cells()][rng.permutation(base.num_cells())]
    mesh = Mesh(X[perm], cells)
    assert mesh.structured is None
    problem, solver, sym = build_problem(3, CFG1, precision, direct, mesh=mesh, **extra)
    ref = oracle_steps(solver, CFG1, sym, 0.05, 1.0, 0.1)
    space = solver.up_.function_space()
    k = 0
    for (up, t) in solver.solve():
        errs = field_errors(space, up.vector().get_local(), ref[k][0])
        print("unstructured", precision, k, errs, flush=True)
        assert max(errs) <= 1e-8, (precision, k, errs)
        k += 1
    assert k == len(ref) == 2
