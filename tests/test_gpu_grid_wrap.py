"""Uniform grid with heavily wrapped cell keys: a context sized for a few
hundred colliders has a table of 2^3k cells with k as small as 2-3 bits per
axis, so bodies spread over a large world share keys with far cells.  The
pair set must still be exactly the reference's (every inclusive AABB overlap,
once, in canonical order), with the half-neighbourhood query finding each
pair across cells once and within a cell by the lower index
(broadphase.cpp:90-182)."""
import numpy as np
import pytest

from physgpu import abi, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,half_side,seed", [(16, 6.0, 1), (200, 12.0, 2), (400, 60.0, 3),
                                              (400, 4.0, 4)])
def test_wrapped_keys_pairs_exact(physgpu, oracle, n, half_side, seed):
    print(f"case n={n} half_side={half_side} seed={seed}", flush=True)
    sc = scenes.spheres(n, seed=seed, half_side=half_side)
    pipe = physgpu.CollisionPipeline(n, 1 << 16)
    try:
        pipe.set_colliders(sc)
        aabbs = pipe.compute_aabbs()
        got = pipe.find_candidate_pairs(algorithm=abi.UNIFORM_GRID)
        want, _ = oracle.find_pairs(aabbs, abi.UNIFORM_GRID)
        assert got.shape == want.shape
        assert np.array_equal(got, want)
        # The same pairs with every 7th body moved onto a neighbour, so that
        # cells hold several bodies and pairs cross every kind of offset.
        pos = sc.pose.copy()
        for i in range(0, n - 1, 7):
            pos[:3, i] = pos[:3, i + 1] + np.float32(0.3)
        pipe.update_poses(pos)
        aabbs = pipe.compute_aabbs()
        got = pipe.find_candidate_pairs(algorithm=abi.UNIFORM_GRID)
        want, _ = oracle.find_pairs(aabbs, abi.UNIFORM_GRID)
        assert len(want) >= n // 7  # the moved bodies' AABBs overlap their neighbours'
        assert np.array_equal(got, want)
    finally:
        pipe.close()
