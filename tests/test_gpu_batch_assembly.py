"""Batched assembly: one launch per colour index assembles that colour of
every instance (mpet_assemble_primitives_batched).  Each instance's
primitives must be bitwise those of its own assembly (same lanes, same
order); batches built on them are checked against the oracle by the batch
tests."""
import pytest
import torch

from test_gpu_batch import INSTANCES, _make

pytestmark = pytest.mark.gpu


def test_batched_assembly_is_bitwise_the_per_instance_assembly():
    from waterscapes_amd.device import assemble_batch
    from waterscapes_amd.solver import batch_system
    items = _make(range(len(INSTANCES)))
    solvers = [it[2] for it in items]

    def systems(assemble):
        return [batch_system(s.problem, s.up_.function_space(), s.params["dt"],
                             s.params["theta"], s.params["device"], assemble=assemble)
                for s in solvers]
    alone = systems(True)
    together = systems(False)
    assert not any(s.assembled for s in together)
    assemble_batch(together)
    torch.cuda.synchronize()
    for a, b in zip(alone, together):
        assert b.assembled
        for name in ("Kuu", "Lap", "Dv", "Dt", "M1", "St", "StK"):
            assert torch.equal(a.prim_t[name], b.prim_t[name]), name
