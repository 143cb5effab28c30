"""The pairs of the 1,048,576-body mixed scene whose EPA polytope outgrows the
GPU's LDS bounds, against the reference's own unbounded computeEpa
(src/collision/epa.cpp:64-130) recorded in tests/golden/epa_overflow.npz by
tests/golden/make_epa_overflow.py (from oracle/_ref, the reference compiled
from its sources, one child process per pair with a time and memory limit).

The oracle in the GPU's mode (exact list within the bounds, continued over
distinct faces with copy counts past them) must give the reference's status,
normal and depth bits on every pair the reference finished; without the
continuation every one of these pairs overflows."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from physgpu import abi, scenes

FIXTURE = os.path.join(GOLDEN, "epa_overflow.npz")


@pytest.fixture(scope="module")
def fixture():
    if not os.path.exists(FIXTURE):
        pytest.skip("tests/golden/epa_overflow.npz not written")
    g = np.load(FIXTURE, allow_pickle=False)
    sc = scenes.Scene(pose=g["pose"], shape_type=g["shape_type"], shape_data=g["shape_data"],
                      hull_range=g["hull_range"], hull_points=g["hull_points"])
    m = len(g["pairs"])
    local = np.stack([np.arange(m) * 2, np.arange(m) * 2 + 1], 1).astype(np.uint32)
    return g, sc, local


def test_fixture_pairs_overflow_the_bounds(oracle, fixture):
    g, sc, local = fixture
    assert len(local) > 50 and int(g["hull_pair_count"]) > 30000
    oracle.mirror_gpu_epa(continuation=False)
    try:
        st, _, _, _ = oracle.gjk_epa(sc, local)
    finally:
        oracle.mirror_gpu_epa(False)
    assert np.all(st == 3)


def test_continuation_equals_reference(oracle, fixture):
    g, sc, local = fixture
    oracle.mirror_gpu_epa()
    try:
        st, nrm, dep, _ = oracle.gjk_epa(sc, local)
    finally:
        oracle.mirror_gpu_epa(False)
    ref = g["ref_status"]
    done = ref >= 0
    # The reference converges on some of these pairs: the bound alone would
    # drop their contacts.
    assert (ref == 1).sum() >= 10 and (ref == 2).sum() >= 40
    assert not np.any(st == 3)  # no pair left unresolved
    status = np.where(st == 1, 1, 2)
    assert np.array_equal(status[done], ref[done])
    conv = ref == 1
    assert np.array_equal(nrm[conv].view(np.uint32), g["ref_normal"][conv].view(np.uint32))
    assert np.array_equal(dep[conv].view(np.uint32), g["ref_depth"][conv].view(np.uint32))
