"""The oracle's floating-point operation count (orc_count_flops, the op-counting
FP32 scalar over generateContact), which the narrow phase's VALU roofline
divides by kernel time.  Pinned by what the arithmetic fixes:

  * the counting scalar computes the float path's values: same touching pairs
    and the same manifolds, bit for bit, as generate_contacts;
  * a hull support map costs one dot product (3 multiplies, 2 adds) per hull
    point and the rest of the code does not depend on the point count: with
    every hull point listed twice (the support still picks the first copy,
    so every value and branch stays the same) a pair's count grows by exactly
    5 N per support evaluation of its N-point hull, i.e. by a multiple of 5 N
    equal to about one per GJK / EPA iteration;
  * counts do not depend on the number of worker threads, and a plain
    sphere-sphere contact has one count for every touching pair (no branch on
    values past the overlap test)."""
import numpy as np
import pytest

from physgpu import abi, scenes


@pytest.fixture(scope="module")
def case(oracle):
    sc = scenes.mixed(8192, seed=71, side=scenes.mixed_side_for(8192, 1.5))
    pairs, _ = oracle.find_pairs(oracle.compute_aabbs(sc), abi.UNIFORM_GRID)
    return sc, pairs


def _doubled_hulls(sc):
    """The scene with every hull's points listed twice (same order, twice)."""
    hr = sc.hull_range.copy()
    pts = []
    cursor = 0
    for i in range(sc.count):
        if sc.shape_type[i] != abi.HULL:
            continue
        f, c = (int(x) for x in sc.hull_range[:, i])
        p = sc.hull_points[:, f:f + c]
        pts.append(np.concatenate([p, p], axis=1))
        hr[:, i] = (cursor, 2 * c)
        cursor += 2 * c
    out = scenes.Scene(pose=sc.pose, shape_type=sc.shape_type, shape_data=sc.shape_data,
                       hull_range=hr, hull_points=np.ascontiguousarray(np.concatenate(pts, axis=1)))
    return out


def test_counting_scalar_is_the_float_path(oracle, case):
    sc, pairs = case
    f = oracle.count_flops(sc, pairs)
    rec, hit = oracle.generate_contacts(sc, pairs)
    assert np.array_equal(f[:, 7].astype(bool), hit)
    assert np.array_equal(oracle.count_flops(sc, pairs, workers=8), f)
    # Sphere-sphere: one count for every touching pair.
    ss = (rec["shape_pair"] == 0) & hit
    assert ss.sum() > 100
    totals = f[ss, :4].sum(axis=1)
    assert len(np.unique(totals)) == 1 and totals[0] > 0
    assert (f[ss, 3] == 1).all()  # one square root (the distance)


def test_hull_point_cost_is_one_dot_product(oracle, case):
    sc, pairs = case
    t = sc.shape_type
    a_hull, b_hull = t[pairs[:, 0]] == abi.HULL, t[pairs[:, 1]] == abi.HULL
    one = a_hull ^ b_hull  # exactly one hull in the pair
    sel = pairs[one]
    assert len(sel) > 200
    dsc = _doubled_hulls(sc)
    f1 = oracle.count_flops(sc, sel)
    f2 = oracle.count_flops(dsc, sel)
    # Same values and branches: same iterations, same touching pairs, and the
    # same manifolds from the float path.
    assert np.array_equal(f1[:, 5:8], f2[:, 5:8])
    r1, h1 = oracle.generate_contacts(sc, sel)
    r2, h2 = oracle.generate_contacts(dsc, sel)
    assert np.array_equal(h1, h2)
    assert np.array_equal(r1.view(np.uint8), r2.view(np.uint8))
    # Only multiplies and adds grow, by 3 N and 2 N per support evaluation.
    assert np.array_equal(f1[:, 2:4], f2[:, 2:4])
    hull = np.where(t[sel[:, 0]] == abi.HULL, sel[:, 0], sel[:, 1])
    n = sc.hull_range[1, hull].astype(np.int64)
    dmul = (f2[:, 1] - f1[:, 1]).astype(np.int64)
    dadd = (f2[:, 0] - f1[:, 0]).astype(np.int64)
    assert (dmul % (3 * n) == 0).all() and (dadd % (2 * n) == 0).all()
    calls = dmul // (3 * n)
    assert np.array_equal(calls, dadd // (2 * n))
    # About one support evaluation of the hull per GJK / EPA iteration (the
    # last iteration of a loop may stop before its support; a contact takes
    # one more for its point).
    iters = (f1[:, 5] + f1[:, 6]).astype(np.int64)
    assert (calls >= iters - 2).all() and (calls <= iters + 1).all()
    assert (calls > 0).all()
