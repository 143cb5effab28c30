"""Pins of the oracle's FP64 narrow-phase adjoint (forward-mode dual numbers
over the oracle's templated contact code, oracle/physcpu.cpp), which the GPU
adjoint is checked against:
  * FP64 central differences (h = 1e-6) of the same linearised loss on
    isolated touching pairs covering every SAT axis, the fallback point,
    clipped points from both faces, the sphere-box face / edge / corner
    regions and the inside branch;
  * closed forms: d(depth)/d(position) = +n for the first collider and -n for
    the second (translating a body along the normal changes the depth at unit
    rate), translation invariance of the normal and of the points' relative
    geometry, sphere-sphere normal derivative (I - n n^T) / |d|, and zero
    gradients for sphere orientations."""
import numpy as np
import pytest

from physgpu import abi, scenes


@pytest.fixture(scope="module")
def touching():
    sc = scenes.contact_pairs(2048, seed=4, inside_fraction=0.25)
    pairs = np.stack([np.arange(0, sc.count, 2), np.arange(1, sc.count, 2)], 1).astype(np.uint32)
    return sc, pairs


def _loss_seeds(oracle, sc, pairs, seed=1):
    normal, depth, point, count = oracle.contact_geometry_f64(sc, pairs)
    m = len(pairs)
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((3, m))
    w /= np.linalg.norm(w, axis=0)
    gd = 2.0 * depth  # d(sum depth^2)
    gp = rng.standard_normal((12, m)) * 0.1
    used = np.repeat(np.arange(4)[:, None] < count[None, :], 3, axis=0)
    return w.astype(np.float32), gd.astype(np.float32), (gp * used).astype(np.float32)


def test_touching_pairs_cover_every_branch(oracle, touching):
    sc, pairs = touching
    ch = oracle.contact_choices(sc, pairs)
    assert ch["hit"].all()
    bb = (sc.shape_type[pairs[:, 0]] == abi.BOX) & (sc.shape_type[pairs[:, 1]] == abi.BOX)
    assert set(np.unique(ch["sel"][bb])) == set(range(15))
    assert ch["fallback"][bb].sum() > 50
    assert ch["from_b"][bb].sum() > 50
    assert (ch["count"][bb] > 1).sum() > 0
    sb = ~bb
    assert set(np.unique(ch["clamped"][sb & (ch["branch"] == 0)])) == set(range(1, 8))
    assert (ch["branch"][sb] == 1).sum() > 50


def test_dual_adjoint_matches_central_differences(oracle, touching):
    sc, pairs = touching
    gn, gd, gp = _loss_seeds(oracle, sc, pairs)
    ga, status = oracle.contact_adjoint_f64(sc, pairs, gn, gd, gp)
    gf, stable = oracle.contact_fd_f64(sc, pairs, gn, gd, gp, h=1e-6)
    assert (status == 1).all()
    assert stable.mean() > 0.99
    scale = np.maximum(np.abs(ga).max(0), 1e-3)
    rel = (np.abs(ga - gf).max(0) / scale)[stable]
    assert rel.max() < 1e-5, rel.max()


def test_depth_translation_gradient_is_the_normal(oracle, touching):
    sc, pairs = touching
    normal, depth, _, count = oracle.contact_geometry_f64(sc, pairs)
    m = len(pairs)
    for k in range(2):
        gd = np.zeros((4, m), np.float32)
        gd[k] = 1.0
        ga, status = oracle.contact_adjoint_f64(sc, pairs, g_depth=gd)
        ch = oracle.contact_choices(sc, pairs)
        live = (count > k) & ((ch["zero_depths"] == 0) | (k > 0))
        np.testing.assert_allclose(ga[0:3, live], normal[:, live], atol=1e-9)
        np.testing.assert_allclose(ga[7:10, live], -normal[:, live], atol=1e-9)


def test_translation_invariance(oracle, touching):
    sc, pairs = touching
    gn, gd, gp = _loss_seeds(oracle, sc, pairs, seed=7)
    ga, _ = oracle.contact_adjoint_f64(sc, pairs, g_normal=gn, g_depth=gd)
    # Normal and depth do not change when both bodies move together.
    np.testing.assert_allclose(ga[0:3] + ga[7:10], 0.0, atol=1e-9)
    # Points move with a common translation: sum of position gradients = sum of point seeds.
    _, _, _, count = oracle.contact_geometry_f64(sc, pairs)
    gpt, _ = oracle.contact_adjoint_f64(sc, pairs, g_point=gp)
    want = gp.reshape(4, 3, -1).sum(0)
    np.testing.assert_allclose(gpt[0:3] + gpt[7:10], want, atol=1e-9)


def test_sphere_sphere_closed_form(oracle):
    rng = np.random.default_rng(5)
    m = 64
    pa = rng.uniform(-1, 1, (3, m)) * 0.1
    d = rng.standard_normal((3, m))
    d /= np.linalg.norm(d, axis=0)
    ra, rb = rng.uniform(0.2, 0.5, m), rng.uniform(0.2, 0.5, m)
    dist = (ra + rb) * rng.uniform(0.5, 0.95, m)
    pos = np.zeros((7, 2 * m), np.float32)
    q = scenes.random_quaternions(rng, 2 * m)
    pos[3:] = q
    pos[:3, 0::2] = pa
    pos[:3, 1::2] = pa + d * dist
    data = np.zeros((3, 2 * m), np.float32)
    data[0, 0::2], data[0, 1::2] = ra, rb
    sc = scenes.Scene(pose=pos, shape_type=np.zeros(2 * m, np.uint8), shape_data=data)
    pairs = np.stack([np.arange(0, 2 * m, 2), np.arange(1, 2 * m, 2)], 1).astype(np.uint32)
    normal, _, _, _ = oracle.contact_geometry_f64(sc, pairs)
    delta = pos[:3, 1::2].astype(np.float64) - pos[:3, 0::2]
    L = np.linalg.norm(delta, axis=0)
    w = rng.standard_normal((3, m)).astype(np.float32)
    ga, _ = oracle.contact_adjoint_f64(sc, pairs, g_normal=w)
    want_b = (w - normal * (normal * w).sum(0)) / L
    np.testing.assert_allclose(ga[7:10], want_b, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(ga[3:7], 0.0, atol=1e-12)
    np.testing.assert_allclose(ga[10:14], 0.0, atol=1e-12)


def test_hull_adjoint_matches_central_differences(oracle):
    """Box-hull and hull-hull contacts through GJK + EPA: the dual-number
    derivative of normal . w + depth^2 equals FP64 central differences where
    the EPA face does not change under the steps (the contact point itself
    jumps between tied support features of a face, so it is left out)."""
    # FP64 runs stop at the LDS bounds (no continuation): the GPU adjoint
    # differentiates only runs inside them, and past them FP64 can settle on a
    # neighbouring near-coplanar face within the same-face normal tolerance.
    oracle.mirror_gpu_epa(continuation=False)
    try:
        sc = scenes.mixed(20000, seed=5, side=scenes.mixed_side_for(20000, 2.0))
        pairs, _ = oracle.find_pairs(oracle.compute_aabbs(sc), abi.UNIFORM_GRID)
        rec, hit = oracle.generate_contacts(sc, pairs)
        m = rec[hit]
        m = m[(m["shape_pair"] == abi.PAIR_BOX_HULL) | (m["shape_pair"] == abi.PAIR_HULL_HULL)]
        P = np.stack([m["first"], m["second"]], 1).astype(np.uint32)
        M = len(P)
        assert M > 500
        rng = np.random.default_rng(1)
        w = rng.standard_normal((3, M))
        w = (w / np.linalg.norm(w, axis=0)).astype(np.float32)
        gd = np.zeros((4, M), np.float32)
        gd[0] = 2 * m["penetration"][:, 0]
        ga, status = oracle.contact_adjoint_f64(sc, P, w, gd)
        gf, stable = oracle.contact_fd_f64(sc, P, w, gd, h=1e-6)
        assert (status == 1).mean() > 0.99
        ok = stable & (status == 1)
        assert ok.mean() > 0.95
        scale = np.maximum(np.abs(ga).max(0), 1e-3)
        assert (np.abs(ga - gf).max(0) / scale)[ok].max() < 1e-5
    finally:
        oracle.mirror_gpu_epa(False)
