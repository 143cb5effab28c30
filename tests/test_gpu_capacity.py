"""Pair capacity overflow on the GPU: every broad phase reports it
(PHYSGPU_CAPACITY_EXCEEDED -> physgpu.CapacityError) with the full count,
writes nothing past the context's pair buffers, and leaves the context usable:
the next step that fits gives the oracle's pairs and manifolds."""
import numpy as np
import pytest

from conftest import assert_manifolds_equal, compact
from physgpu import abi, scenes

pytestmark = pytest.mark.gpu

ALGOS = [abi.SWEEP_AND_PRUNE, abi.UNIFORM_GRID, abi.DYNAMIC_TREE]


@pytest.mark.parametrize("algo", ALGOS)
def test_pair_overflow_reported_then_recovers(physgpu, oracle, algo):
    big = scenes.spheres(4096, seed=11)            # about 3,500 pairs
    small = scenes.spheres(1024, seed=12, half_side=10.0)
    want_small, _ = oracle.find_pairs(oracle.compute_aabbs(small), abi.SWEEP_AND_PRUNE)
    assert 0 < len(want_small) < 1024
    with physgpu.CollisionPipeline(4096, 1024) as p:
        print(f"overflow algo {algo}", flush=True)
        p.set_colliders(big)
        with pytest.raises(physgpu.CapacityError):
            p.collide(algorithm=algo)
        with pytest.raises(physgpu.CapacityError):
            p.find_candidate_pairs(algorithm=algo)
        # The context still works for a step that fits.
        p.set_colliders(small)
        npairs, nman = p.collide(algorithm=algo)
        assert npairs == len(want_small)
        assert np.array_equal(p.pairs(), want_small)
        rec, hit = oracle.generate_contacts(small, want_small)
        assert_manifolds_equal(p.manifolds(), compact(rec, hit), "after overflow")
