"""Full-size checks on the 48^3 x 6-tet, A = 4 benchmark workload (3,326,264
DOFs), where the oracle is too slow: size-independent properties only.

* the assembled operator is symmetric: x.(A y) = y.(A x) for random x, y, in
  the FP64 structured form, the FP64 scalar CSR and the FP32 copy; the
  structured FP64 apply equals the CSR apply;
* the block preconditioner is symmetric (r1.(M r2) = r2.(M r1)) in FP64;
* one mixed-precision step (FP32 MINRES/AMG in FP64 refinement) equals a
  tightly converged FP64 step to 1e-8 relative L2 per field;
* repeating a step from the same state is bitwise reproducible."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def workload():
    sys.path.insert(0, ROOT)
    import bench
    mixed = bench.make_solver(48, 4, "mixed")
    ref = bench.make_solver(48, 4, "fp64", krylov_rtol=1e-13)
    return mixed, ref


def _rel(a, b):
    return float(abs(a - b) / max(abs(a), abs(b), 1e-300))


def test_full_size_operator_symmetry(workload):
    import torch
    mixed, ref = workload
    s = ref.engine().system
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.randn(s.n, device="cuda", dtype=torch.float64, generator=g)
    y = torch.randn(s.n, device="cuda", dtype=torch.float64, generator=g)
    Ax, Ay = s.apply(x, masked=False), s.apply(y, masked=False)
    scale = float(torch.linalg.norm(Ax) * torch.linalg.norm(y))
    assert abs(float(x @ Ay - y @ Ax)) <= 1e-12 * scale
    Cx = s.csr_apply(x, masked=False)
    assert float(torch.linalg.norm(Cx - Ax)) <= 1e-12 * float(torch.linalg.norm(Ax))
    x32, y32 = x.float(), y.float()
    Ax32 = mixed.engine().system.apply(x32, dtype=1, masked=False).double()
    Ay32 = mixed.engine().system.apply(y32, dtype=1, masked=False).double()
    assert abs(float(x32.double() @ Ay32 - y32.double() @ Ax32)) <= 1e-5 * scale
    assert float(torch.linalg.norm(Ax32 - Ax)) <= 1e-5 * float(torch.linalg.norm(Ax))


def test_full_size_preconditioner_symmetry(workload):
    import torch
    mixed, ref = workload
    kr = ref.engine().krylov
    s = ref.engine().system
    g = torch.Generator(device="cuda").manual_seed(5)
    r1 = torch.randn(s.n, device="cuda", dtype=torch.float64, generator=g)
    r2 = torch.randn(s.n, device="cuda", dtype=torch.float64, generator=g)
    free = (s.dmask == 0).double()
    r1, r2 = r1 * free, r2 * free
    z1, z2 = kr.precond(r1, 0), kr.precond(r2, 0)
    a, b = float(r2 @ z1), float(r1 @ z2)
    assert _rel(a, b) <= 1e-10, (a, b)
    assert float(r1 @ z1) > 0 and float(r2 @ z2) > 0


def test_full_size_mixed_step_matches_fp64(workload):
    import torch
    mixed, ref = workload
    space = ref.up_.function_space()
    gm, gr = mixed.solve(), ref.solve()
    for k in range(2):
        upm, tm = next(gm)
        upr, tr = next(gr)
        assert tm == tr
        a, b = upm.data, upr.data
        for i in range(space.num_sub_spaces()):
            idx = torch.as_tensor(space.field_index(i), device="cuda")
            e = float(torch.linalg.norm(a[idx] - b[idx]) / torch.linalg.norm(b[idx]))
            assert e <= 1e-8, (k, i, e)


def test_full_size_step_is_reproducible(workload):
    import torch
    mixed, ref = workload
    eng = mixed.engine()
    up_prev = mixed.up_.data.clone()
    t = float(mixed.problem.time)
    x1 = torch.empty_like(up_prev)
    x2 = torch.empty_like(up_prev)
    eng.hist = []
    eng.step(up_prev, x1, t)
    mixed.problem.time.assign(t)
    eng.hist = []
    eng.step(up_prev, x2, t)
    mixed.problem.time.assign(t)
    torch.cuda.synchronize()
    assert torch.equal(x1, x2)
