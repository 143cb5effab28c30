"""Pins of the oracle's hull contact (GJK + EPA manifolds), which the reference
cannot pin: it has no hull shape (include/phys/collision/shapes/convex_hull.hpp:5)
and no caller of computeEpa (src/collision/epa.cpp:64-130).

Every hull manifold of a dense mixed scene is re-derived by brute force in
float64 (tests/hullcheck.py): its contact point must be the midpoint of the
two deepest features along the EPA normal, both anchors must map back to that
point, and its depth must be the Minkowski support distance along the normal.
A planted sign or weight mistake in the oracle's contact point must fail it."""
import numpy as np
import pytest

import hullcheck
from physgpu import abi, scenes


@pytest.fixture(autouse=True)
def bounded_epa(oracle):
    # The GPU's polytope bounds keep every EPA call short (a few FP32 pairs
    # grow the reference's unbounded polytope for minutes); the contact point
    # of a converged pair does not depend on them.
    oracle.mirror_gpu_epa()
    yield
    oracle.mirror_gpu_epa(False)


@pytest.fixture(scope="module")
def hull_scene(oracle):
    n = 6000
    sc = scenes.mixed(n, seed=31, side=scenes.mixed_side_for(n, 2.5))
    aabbs = oracle.compute_aabbs(sc)
    pairs, _ = oracle.find_pairs(aabbs, abi.UNIFORM_GRID)
    hull = sc.shape_type == abi.HULL
    hp = pairs[hull[pairs[:, 0]] | hull[pairs[:, 1]]]
    return sc, hp


def _contacts(oracle, sc, pairs):
    rec, hit = oracle.generate_contacts(sc, pairs)
    return rec[hit]


def test_hull_contact_point_is_support_midpoint(oracle, hull_scene):
    sc, hp = hull_scene
    man = _contacts(oracle, sc, hp)
    classes = np.bincount(man["shape_pair"], minlength=6)
    # Every hull class is exercised.
    assert classes[abi.PAIR_SPHERE_HULL] > 50 and classes[abi.PAIR_BOX_HULL] > 50
    assert classes[abi.PAIR_HULL_HULL] > 10
    r = hullcheck.check(sc, man)
    assert len(r["index"]) == len(man)
    assert r["anchors"].all(), np.nonzero(~r["anchors"])[0][:10]
    assert r["midpoint"].all(), np.nonzero(~r["midpoint"])[0][:10]
    assert r["depth"].all(), np.nonzero(~r["depth"])[0][:10]


@pytest.mark.parametrize("mutation", [1, 2])
def test_planted_contact_point_mistakes_fail(oracle, hull_scene, mutation):
    """1: supports taken along the wrong signs; 2: A's support point instead of
    the midpoint.  Both must fail the brute-force check on most manifolds
    (the midpoint and A's support coincide only on touching contacts)."""
    sc, hp = hull_scene
    oracle.set_test_mutation(mutation)
    try:
        man = _contacts(oracle, sc, hp)
    finally:
        oracle.set_test_mutation(0)
    r = hullcheck.check(sc, man)
    assert r["anchors"].all()  # anchors still map to the (wrong) point
    assert (~r["midpoint"]).mean() > 0.9, (~r["midpoint"]).mean()


def test_hull_contact_in_fp64_matches_brute_force(oracle, hull_scene):
    """The FP64 re-evaluation's hull contact points and depths agree with the
    FP32 manifolds to FP32 rounding where both converge."""
    sc, hp = hull_scene
    rec, hit = oracle.generate_contacts(sc, hp)
    d, hit64 = oracle.generate_contacts_f64(sc, hp)
    both = hit & hit64
    assert both.sum() > 0.97 * hit.sum()
    nerr = np.abs(rec["normal"][both] - d[both, 1:4]).max(axis=1)
    # EPA in FP32 and FP64 converge on the same face on nearly every pair.
    assert (nerr < 1e-3).mean() > 0.95
