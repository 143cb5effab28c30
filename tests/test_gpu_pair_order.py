"""Canonical pair order on the GPU for every segment size the segmented sort
distinguishes (pairs per `first` body: 2..16 in registers, 17..64 by a wave,
65..4096 bitonic in LDS, beyond that LDS bitmap windows, including one body
with a pair against each of 300,014 others, whose keys span two windows).

Expected pairs come from the CPU oracle's broad phase (the reference's
lexicographic order, src/collision/broadphase.cpp:18-60); the scene fixes the
count per segment independently of any sort (scenes.pair_segments)."""
import numpy as np
import pytest

from physgpu import abi, scenes

pytestmark = pytest.mark.gpu

ALGOS = [abi.SWEEP_AND_PRUNE, abi.UNIFORM_GRID, abi.DYNAMIC_TREE]


@pytest.fixture(scope="module")
def pipe(physgpu):
    p = physgpu.CollisionPipeline(310000, 1 << 22)
    yield p
    p.close()


@pytest.fixture(scope="module")
def segments_case(oracle):
    sc = scenes.pair_segments()
    want, _ = oracle.find_pairs(oracle.compute_aabbs(sc), abi.SWEEP_AND_PRUNE)
    return sc, want


@pytest.fixture(scope="module")
def dense_case(oracle):
    # About 80 pairs per body as `first`: mostly wave and LDS segments.
    sc = scenes.spheres(20000, seed=31, half_side=3.0)
    want, _ = oracle.find_pairs(oracle.compute_aabbs(sc), abi.SWEEP_AND_PRUNE)
    return sc, want


def _lexicographic(pairs):
    key = (pairs[:, 0].astype(np.int64) << 32) | pairs[:, 1].astype(np.int64)
    return bool(np.all(np.diff(key) > 0))


@pytest.mark.parametrize("algo", ALGOS)
def test_segment_sizes(pipe, segments_case, algo):
    sc, want = segments_case
    print(f"pair_segments n={sc.count} algo {algo}", flush=True)
    pipe.set_colliders(sc)
    npairs, _ = pipe.collide(algorithm=algo)
    got = pipe.pairs()
    nb = len(sc.meta["segments"])
    assert list(np.bincount(want[:, 0], minlength=nb)[:nb]) == sc.meta["segments"]
    assert npairs == len(want)
    assert _lexicographic(got)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("algo", ALGOS)
def test_dense_segments(pipe, dense_case, algo):
    sc, want = dense_case
    print(f"dense spheres n={sc.count} algo {algo}", flush=True)
    pipe.set_colliders(sc)
    npairs, _ = pipe.collide(algorithm=algo)
    got = pipe.pairs()
    per_first = np.bincount(want[:, 0])
    assert per_first.max() > 64 and (per_first > 16).sum() > 1000
    assert npairs == len(want)
    assert np.array_equal(got, want)


def test_repeated_steps_same_order(pipe, dense_case):
    """Segment counters and lists are cleared on the device every step."""
    sc, want = dense_case
    pipe.set_colliders(sc)
    for algo in (abi.UNIFORM_GRID, abi.UNIFORM_GRID, abi.DYNAMIC_TREE):
        pipe.collide(algorithm=algo)
        assert np.array_equal(pipe.pairs(), want)
