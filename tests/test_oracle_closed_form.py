"""Pins of the oracle that do not come from the reference binary: brute force,
the reference's known-answer tests (tests/collision/narrowphase_tests.cpp),
closed forms and invariants.  Each is chosen so a plausible slip (wrong sign,
swapped operand, dropped term, wrong index) fails."""
import numpy as np
import pytest

from conftest import golden
from physgpu import abi, scenes


def brute_pairs(b):
    b = np.asarray(b, np.float32)
    n = len(b)
    i, j = np.triu_indices(n, 1)
    ov = ((b[i, 0] <= b[j, 3]) & (b[i, 3] >= b[j, 0]) & (b[i, 1] <= b[j, 4]) &
          (b[i, 4] >= b[j, 1]) & (b[i, 2] <= b[j, 5]) & (b[i, 5] >= b[j, 2]))
    return np.stack([i[ov], j[ov]], 1).astype(np.uint32)


@pytest.mark.parametrize("algo", [0, 1, 2])
def test_pairs_equal_brute_force(oracle, algo):
    for seed in range(3):
        sc = scenes.mixed(1500, seed=100 + seed, side=scenes.mixed_side_for(1500, 2.0))
        b = oracle.compute_aabbs(sc)
        pairs, st = oracle.find_pairs(b, algo)
        assert np.array_equal(pairs, brute_pairs(b))
        assert st["aabb_pairs"] == len(pairs)
        if algo == 0:
            # xWindowComparisons = number of x-overlapping pairs (broadphase_tests.cpp:22-24)
            i, j = np.triu_indices(len(b), 1)
            xo = ((b[i, 0] <= b[j, 3]) & (b[i, 3] >= b[j, 0])).sum()
            assert st["x_window_comparisons"] == xo


def two(pa, sa, pb, sb):
    pose = np.zeros((7, 2), np.float32)
    pose[:, 0] = pa
    pose[:, 1] = pb
    types = np.array([sa[0], sb[0]], np.uint8)
    data = np.zeros((3, 2), np.float32)
    data[:, 0] = sa[1]
    data[:, 1] = sb[1]
    return scenes.Scene(pose=pose, shape_type=types, shape_data=data)


def world(anchor, pose):
    """transform(v, tf) = rotate(q, v) + p in float64 (for tolerance checks)."""
    x, y, z, w = [float(c) for c in pose[3:]]
    q = np.array([w, x, y, z])
    v = np.array([0.0, *anchor])

    def mul(a, b):
        return np.array([a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3],
                         a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2],
                         a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1],
                         a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0]])
    r = mul(mul(q, v), q * np.array([1, -1, -1, -1]))
    return r[1:] + np.asarray(pose[:3], np.float64)


def qz(angle):
    return [0, 0, np.sin(angle / 2), np.cos(angle / 2)]


def test_kat_sphere_sphere(oracle):
    """narrowphase_tests.cpp:35-59: anchors (+-0.75, 0, 0)."""
    sc = two([0, 0, 0, 0, 0, 0, 1], (0, [1, 0, 0]), [1.5, 0, 0, 0, 0, 0, 1], (0, [1, 0, 0]))
    m, hit = oracle.generate_contacts(sc, [[0, 1]])
    assert hit[0] and m[0]["point_count"] == 1
    np.testing.assert_allclose(m[0]["anchor_a"][0], [0.75, 0, 0], atol=1e-5)
    np.testing.assert_allclose(m[0]["anchor_b"][0], [-0.75, 0, 0], atol=1e-5)
    np.testing.assert_allclose(m[0]["normal"], [1, 0, 0], atol=1e-6)
    assert abs(m[0]["penetration"][0] - 0.5) <= 1e-6


def test_kat_box_sphere_order(oracle):
    """narrowphase_tests.cpp:61-86: box (rotated pi/2 about z) then sphere."""
    sc = two([0, 0, 0] + qz(1.57079632679), (1, [1, 1, 1]), [0, 1.5, 0, 0, 0, 0, 1],
             (0, [1, 0, 0]))
    m, hit = oracle.generate_contacts(sc, [[0, 1]])
    assert hit[0]
    np.testing.assert_allclose(m[0]["anchor_a"][0], [0.75, 0, 0], atol=1e-5)
    np.testing.assert_allclose(m[0]["anchor_b"][0], [0, -0.75, 0], atol=1e-5)
    np.testing.assert_allclose(world(m[0]["anchor_a"][0], sc.pose[:, 0]), [0, 0.75, 0], atol=1e-5)
    np.testing.assert_allclose(m[0]["normal"], [0, 1, 0], atol=1e-5)


@pytest.mark.parametrize("angle,height,reverse", [
    (0.3, 0.5 * (np.sin(0.3) + np.cos(0.3)) - 0.02, False),
    (0.3, 0.5 * (np.sin(0.3) + np.cos(0.3)) - 0.02, True),
    (0.1, 0.3, False), (0.1, 0.3, True)])
def test_kat_tilted_box_on_floor(oracle, angle, height, reverse):
    """narrowphase_tests.cpp:109-166: penetrating corners, depth = -vertex.y,
    contact = (x, y/2, z), normal (0, +-1, 0)."""
    floor = ([0, -0.5, 0, 0, 0, 0, 1], (1, [9, 0.5, 4]))
    box = ([0, height, 0] + qz(angle), (1, [0.5, 0.5, 0.5]))
    a, b = (box, floor) if reverse else (floor, box)
    sc = two(a[0], a[1], b[0], b[1])
    m, hit = oracle.generate_contacts(sc, [[0, 1]])
    assert hit[0]
    bp = np.asarray(box[0], np.float32)
    corners = [world([x, -0.5, z], bp) for x in (-0.5, 0.5) for z in (-0.5, 0.5)]
    corners = [c for c in corners if c[1] <= 0.0]
    assert m[0]["point_count"] == len(corners)
    np.testing.assert_allclose(m[0]["normal"], [0, -1 if reverse else 1, 0], atol=1e-5)
    for c in corners:
        want = np.array([c[0], c[1] * 0.5, c[2]])
        ok = False
        for k in range(m[0]["point_count"]):
            wa = world(m[0]["anchor_a"][k], sc.pose[:, 0])
            wb = world(m[0]["anchor_b"][k], sc.pose[:, 1])
            if (np.abs(wa - want).max() <= 1e-5 and np.abs(wb - want).max() <= 1e-5
                    and abs(m[0]["penetration"][k] + c[1]) <= 1e-5):
                ok = True
        assert ok


def test_box_box_normal_unit(oracle):
    """narrowphase_tests.cpp:88-107 and the same on random pairs."""
    sc = scenes.mixed(3000, seed=5, side=scenes.mixed_side_for(3000, 2.0), with_hulls=False)
    pairs, _ = oracle.find_pairs(oracle.compute_aabbs(sc))
    m, hit = oracle.generate_contacts(sc, pairs)
    bb = hit & (m["shape_pair"] == abi.PAIR_BOX_BOX)
    assert bb.sum() > 50
    n = np.linalg.norm(m["normal"][bb].astype(np.float64), axis=1)
    assert np.abs(n - 1).max() <= 1e-5


def test_sphere_sphere_closed_form(oracle):
    """Depth r1 + r2 - d and normal (pB - pA)/d, to FP32 rounding."""
    sc = scenes.spheres(3000, seed=3)
    pairs, _ = oracle.find_pairs(oracle.compute_aabbs(sc))
    m, hit = oracle.generate_contacts(sc, pairs)
    p = sc.pose[:3].T.astype(np.float64)
    r = sc.shape_data[0].astype(np.float64)
    a, b = pairs[:, 0], pairs[:, 1]
    d = np.linalg.norm(p[b] - p[a], axis=1)
    assert np.array_equal(hit, d < r[a] + r[b])
    np.testing.assert_allclose(m["penetration"][hit, 0], (r[a] + r[b] - d)[hit], atol=2e-6)
    np.testing.assert_allclose(m["normal"][hit], ((p[b] - p[a]) / d[:, None])[hit], atol=2e-6)


def test_gjk_epa_sphere_closed_form(oracle):
    """GJK/EPA on sphere pairs: depth -> r1 + r2 - d, normal -> (pB - pA)/d.
    The sphere support map is smooth, so EPA converges to within its
    tolerance plus the polytope's faceting error."""
    sc = scenes.spheres(2000, seed=4)
    pairs, _ = oracle.find_pairs(oracle.compute_aabbs(sc))
    st, nrm, dep, _ = oracle.gjk_epa(sc, pairs, 64, 64, 1e-6)
    p = sc.pose[:3].T.astype(np.float64)
    r = sc.shape_data[0].astype(np.float64)
    a, b = pairs[:, 0], pairs[:, 1]
    d = np.linalg.norm(p[b] - p[a], axis=1)
    depth = r[a] + r[b] - d
    ok = st == 1
    assert ok.sum() > 0.6 * (depth > 1e-3).sum()
    # The reference's EPA (reproduced bit for bit, see the golden test) ends
    # on a wrong face for a few pairs in FP32; the rest sit at the closed form.
    err = np.abs(dep[ok] - depth[ok])
    assert (err < 2e-5).mean() > 0.98
    want_n = (p[b] - p[a]) / d[:, None]
    nerr = np.abs(nrm[ok] - want_n[ok]).max(axis=1)
    assert (nerr < 1e-2).mean() > 0.98
    # GJK never reports a clearly separated pair as intersecting.
    assert not np.any((st > 0) & (depth < -1e-4))


def test_hull_of_box_corners_equals_box(oracle):
    """A hull whose points are the 8 corners of a box has the box's support map
    in every direction without near-zero local components (there the box picks
    the + corner by sign, the hull the first of two equal dot products), so
    GJK/EPA agree bit for bit on almost every pair."""
    rng = np.random.default_rng(9)
    n = 400
    sc = scenes.mixed(n, seed=21, side=scenes.mixed_side_for(n, 4.0), with_hulls=False)
    boxes = np.nonzero(sc.shape_type == abi.BOX)[0]
    corners = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)],
                       np.float32)
    hull = scenes.Scene(pose=sc.pose.copy(), shape_type=sc.shape_type.copy(),
                        shape_data=sc.shape_data.copy())
    hr = np.zeros((2, n), np.uint32)
    pts = []
    for k, i in enumerate(boxes):
        hull.shape_type[i] = abi.HULL
        hr[:, i] = (8 * k, 8)
        pts.append(corners * sc.shape_data[:, i])
    hull.hull_range = hr
    hull.hull_points = np.ascontiguousarray(np.concatenate(pts).T.astype(np.float32))
    a1 = oracle.compute_aabbs(sc)
    a2 = oracle.compute_aabbs(hull)
    # Box AABB uses |R| h (a different formula): equal to rounding only.
    np.testing.assert_allclose(a1, a2, atol=2e-6)
    pairs, _ = oracle.find_pairs(a1)
    s1 = oracle.gjk_epa(sc, pairs)
    s2 = oracle.gjk_epa(hull, pairs)
    same = (s1[0] == s2[0]) & (s1[2].view(np.uint32) == s2[2].view(np.uint32))
    assert same.mean() > 0.95, same.mean()
    assert (s1[0] == 1).sum() > 20
    del rng


def test_fp64_twin_close(oracle):
    """The FP64 re-evaluation agrees with FP32 to ~1e-5 relative on a mixed scene
    where both take the same discrete branches."""
    sc = scenes.mixed(3000, seed=8, side=scenes.mixed_side_for(3000, 1.5), with_hulls=False)
    pairs, _ = oracle.find_pairs(oracle.compute_aabbs(sc))
    m, hit = oracle.generate_contacts(sc, pairs)
    d, hit64 = oracle.generate_contacts_f64(sc, pairs)
    same = hit & hit64 & (m["point_count"] == d[:, 0])
    assert same.sum() > 0.95 * hit.sum()
    err = np.abs(m["normal"][same] - d[same, 1:4]).max()
    assert err < 1e-4


def test_golden_kat_matches_closed_form():
    """The reference's own KAT outputs in the fixture satisfy the KAT bounds."""
    g = golden("kat.npz")
    c = g["contacts"]
    np.testing.assert_allclose(c[0]["anchor_a"][0], [0.75, 0, 0], atol=1e-5)
    np.testing.assert_allclose(c[1]["anchor_b"][0], [0, -0.75, 0], atol=1e-5)
