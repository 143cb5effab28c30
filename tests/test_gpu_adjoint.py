"""The narrow-phase adjoint kernel (libphysgpu.so physgpu_contact_adjoint)
against the oracle's FP64 dual-number adjoint and FP64 central differences
(h = 1e-6), relative error <= 1e-3, on 65,536 isolated touching pairs (half
box-box, half sphere-box, depth U[1e-3, 5e-2] m, away from SAT ties) and on a
batch of 256 worlds x 1,024 mixed bodies.  Loss = sum depth^2 + sum normal . w
(w seeded unit vectors), plus random seeds on the contact points."""
import numpy as np
import pytest

from physgpu import abi, scenes

pytestmark = pytest.mark.gpu

TOL = 1e-3


def _pairs_of(man):
    return np.stack([man["first"], man["second"]], 1).astype(np.uint32)


def _seeds(man, seed=1, points=True):
    m = len(man)
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((3, m))
    w /= np.linalg.norm(w, axis=0)
    gd = np.zeros((4, m), np.float32)
    used = np.arange(4)[:, None] < man["point_count"][None, :]
    gd[:] = np.where(used, 2.0 * man["penetration"].T, 0.0)
    gp = None
    if points:
        gp = (rng.standard_normal((12, m)) * 0.1 * np.repeat(used, 3, axis=0)).astype(np.float32)
    return w.astype(np.float32), gd, gp


def _sphere_separation(sc, man):
    """|D| of sphere-sphere (centre distance) and outside-branch sphere-box
    (radius - depth, the centre's distance to the box) manifolds, in float64;
    inf for the others."""
    a, b = man["first"].astype(np.int64), man["second"].astype(np.int64)
    ta, tb = sc.shape_type[a], sc.shape_type[b]
    out = np.full(len(man), np.inf)
    ss = (ta == abi.SPHERE) & (tb == abi.SPHERE)
    d = sc.pose[:3, b].astype(np.float64) - sc.pose[:3, a].astype(np.float64)
    out[ss] = np.linalg.norm(d[:, ss], axis=0)
    sb = (ta == abi.SPHERE) ^ (tb == abi.SPHERE)
    sb &= (ta != abi.HULL) & (tb != abi.HULL)
    r = np.where(ta == abi.SPHERE, sc.shape_data[0, a], sc.shape_data[0, b]).astype(np.float64)
    gap = r - man["penetration"][:, 0].astype(np.float64)
    out[sb & (gap > 0)] = gap[sb & (gap > 0)]
    return np.maximum(out, 1e-12)


def _rel(got, want):
    scale = np.maximum(np.abs(want).max(0), 1e-3)
    return np.abs(got - want).max(0) / scale


@pytest.fixture(scope="module")
def touching(physgpu):
    sc = scenes.contact_pairs(65536, seed=20261203)
    pipe = physgpu.CollisionPipeline(sc.count, 4 * sc.count)
    pipe.set_colliders(sc)
    npairs, nman = pipe.collide(algorithm=abi.UNIFORM_GRID)
    yield sc, pipe, pipe.manifolds()
    pipe.close()


def test_touching_pairs_all_collide(touching):
    sc, pipe, man = touching
    assert len(man) == 65536
    assert np.array_equal(man["first"], np.arange(0, sc.count, 2))


def test_recomputed_forward_is_bitwise(touching):
    """The adjoint's forward recompute takes the narrow phase's choices: its
    normals and depths equal the manifolds' bits, its points their anchors."""
    sc, pipe, man = touching
    normal, depth, point = pipe.contact_geometry()
    assert np.array_equal(normal.T.view(np.uint32), man["normal"].view(np.uint32))
    assert np.array_equal(depth.T.view(np.uint32), man["penetration"].view(np.uint32))
    # World points vs anchor_a taken to world (FP32 rounding only).
    q = sc.pose[3:, man["first"]].astype(np.float64)
    p = sc.pose[:3, man["first"]].astype(np.float64)
    for k in range(4):
        used = man["point_count"] > k
        a = man["anchor_a"][:, k, :].T.astype(np.float64)
        w = p + scenes._rotate(q, a)
        np.testing.assert_allclose(point[3 * k:3 * k + 3][:, used], w[:, used], atol=2e-5)


def test_adjoint_matches_fp64_dual_and_differences(touching, oracle):
    sc, pipe, man = touching
    gn, gd, gp = _seeds(man)
    got, skipped = pipe.contact_adjoint(gn, gd, gp, per_pair=True)
    assert skipped == 0
    pairs = _pairs_of(man)
    want, status = oracle.contact_adjoint_f64(sc, pairs, gn, gd, gp, workers=16)
    assert (status == 1).all()
    rel = _rel(got.astype(np.float64), want)
    print(f"GPU vs FP64 dual: max rel {rel.max():.2e}, p99.9 {np.percentile(rel, 99.9):.2e}")
    assert rel.max() <= TOL
    fd, stable = oracle.contact_fd_f64(sc, pairs, gn, gd, gp, h=1e-6, workers=16)
    assert stable.mean() > 0.99
    rel_fd = _rel(got.astype(np.float64), fd)[stable]
    print(f"GPU vs FP64 central differences: max rel {rel_fd.max():.2e} on {stable.sum()} pairs")
    assert rel_fd.max() <= TOL


def test_pose_gradients_sum_pair_gradients(touching):
    sc, pipe, man = touching
    gn, gd, gp = _seeds(man, seed=2)
    per, _ = pipe.contact_adjoint(gn, gd, gp, per_pair=True)
    pose, _ = pipe.contact_adjoint(gn, gd, gp)
    want = np.zeros((7, sc.count), np.float64)
    np.add.at(want.T, man["first"], per[:7].T.astype(np.float64))
    np.add.at(want.T, man["second"], per[7:].T.astype(np.float64))
    np.testing.assert_allclose(pose, want, rtol=1e-5, atol=1e-6)


def test_quaternion_gradients_are_tangent(touching):
    sc, pipe, man = touching
    gn, gd, gp = _seeds(man, seed=3)
    per, _ = pipe.contact_adjoint(gn, gd, gp, per_pair=True)
    qa = sc.pose[3:, man["first"]]
    qb = sc.pose[3:, man["second"]]
    assert np.abs((per[3:7] * qa).sum(0)).max() < 1e-4
    assert np.abs((per[10:14] * qb).sum(0)).max() < 1e-4


def test_inside_spheres_and_small_scene(physgpu, oracle):
    sc = scenes.contact_pairs(4096, seed=9, inside_fraction=0.3)
    with physgpu.CollisionPipeline(sc.count, 4 * sc.count) as pipe:
        pipe.set_colliders(sc)
        pipe.collide(algorithm=abi.SWEEP_AND_PRUNE)
        man = pipe.manifolds()
        gn, gd, gp = _seeds(man, seed=4)
        got, _ = pipe.contact_adjoint(gn, gd, gp, per_pair=True)
    want, status = oracle.contact_adjoint_f64(sc, _pairs_of(man), gn, gd, gp)
    ch = oracle.contact_choices(sc, _pairs_of(man))
    assert (ch["branch"] == 1).sum() > 100
    assert _rel(got.astype(np.float64), want).max() <= TOL


def test_batched_mixed_worlds(physgpu, oracle):
    """256 worlds x 1,024 mixed bodies in one batch; sphere-hull manifolds are
    skipped, box-hull and hull-hull go through the EPA face."""
    parts = [scenes.mixed(1024, seed=[77, w], side=scenes.mixed_side_for(1024, 1.33))
             for w in range(256)]
    sc, off = scenes.concat(parts)
    with physgpu.CollisionPipeline(sc.count, 4 * sc.count,
                                   max_hull_points=sc.hull_points.shape[1]) as pipe:
        pipe.set_colliders(sc)
        pipe.set_worlds(off)
        pipe.collide(algorithm=abi.UNIFORM_GRID)
        man = pipe.manifolds()
        gn, gd, gp = _seeds(man, seed=5)
        hull = man["shape_pair"] >= abi.PAIR_SPHERE_HULL
        gp[:, hull] = 0.0  # hull contact points sit on tied supports (see the EPA-face test)
        with pytest.warns(physgpu.AdjointSkippedWarning):
            got, skipped = pipe.contact_adjoint(gn, gd, gp, per_pair=True)
    sphere_hull = man["shape_pair"] == abi.PAIR_SPHERE_HULL
    assert skipped == int(sphere_hull.sum()) > 0
    assert np.all(got[:, sphere_hull] == 0)
    pairs = _pairs_of(man)
    # FP64 runs stop at the LDS bounds (no continuation): the GPU adjoint
    # differentiates only runs inside them, and past them FP64 can settle on a
    # neighbouring near-coplanar face within the same-face normal tolerance.
    oracle.mirror_gpu_epa(continuation=False)
    try:
        want, status = oracle.contact_adjoint_f64(sc, pairs, gn, gd, gp, workers=16)
        fd, stable = oracle.contact_fd_f64(sc, pairs, gn, gd, gp, workers=16)
        n64, _, _, _ = oracle.contact_geometry_f64(sc, pairs)
    finally:
        oracle.mirror_gpu_epa(False)
    assert np.array_equal(status == -1, sphere_hull)
    same_face = np.abs(n64.T - man["normal"]).max(1) < 1e-5
    ok = (status == 1) & (~hull | same_face)
    rel = _rel(got[:, ok].astype(np.float64), want[:, ok])
    # Random scenes are not kept away from ill-conditioned contacts either: a
    # sphere's normal is a unit vector along a separation |D| (centre to
    # centre, or centre to the box's closest point), whose derivative scales
    # as 1/|D|, so FP32 rounding of the ~10 m coordinates (ulp ~1e-6) gives
    # relative errors ~1e-7 * 16 m / |D| there (a centre 5 um off a box face
    # is 3e-3).  The tolerance takes that bound on top of TOL.
    delta = _sphere_separation(sc, man)[ok]
    tol = TOL + 1e-7 * 16.0 / delta
    print(f"batched: {ok.sum()} differentiated manifolds ({(ok & hull).sum()} through EPA), "
          f"max rel {rel.max():.2e}, {(rel > TOL).sum()} above {TOL} "
          f"({(delta < 1e-3).sum()} with |D| < 1 mm); FD-stable {stable[ok].mean():.4f}")
    # Compare where FP64 differences see one smooth branch.
    assert (rel[stable[ok]] <= tol[stable[ok]]).all()
    assert (rel[stable[ok] & (delta >= 1e-3)] <= TOL).all()
    assert ok[hull & ~sphere_hull].mean() > 0.8


def test_torch_autograd_binding(physgpu, oracle):
    torch = pytest.importorskip("torch")
    from physgpu.autograd import contact_geometry
    sc = scenes.contact_pairs(2048, seed=11, inside_fraction=0.2)
    with physgpu.CollisionPipeline(sc.count, 4 * sc.count) as pipe:
        pipe.set_colliders(sc)
        pose = torch.tensor(sc.pose, device="cuda", requires_grad=True)
        normal, depth, point = contact_geometry(pose, pipe)
        m = normal.shape[1]
        rng = np.random.default_rng(6)
        w = rng.standard_normal((3, m))
        w /= np.linalg.norm(w, axis=0)
        wt = torch.tensor(w, device="cuda", dtype=torch.float32)
        loss = (depth ** 2).sum() + (normal * wt).sum()
        loss.backward()
        grad = pose.grad.detach().cpu().numpy()
        man = pipe.manifolds()
    gn = w.astype(np.float32)
    used = np.arange(4)[:, None] < man["point_count"][None, :]
    gd = np.where(used, 2.0 * man["penetration"].T, 0.0).astype(np.float32)
    per, _ = oracle.contact_adjoint_f64(sc, _pairs_of(man), gn, gd)
    want = np.zeros((7, sc.count))
    np.add.at(want.T, man["first"], per[:7].T)
    np.add.at(want.T, man["second"], per[7:].T)
    scale = np.maximum(np.abs(want).max(0), 1e-3)
    assert (np.abs(grad - want).max(0) / scale).max() <= TOL


def test_hull_pairs_through_the_epa_face(physgpu, oracle):
    """Box-hull and hull-hull manifolds: the adjoint recomputes GJK + EPA and
    differentiates the converged face.  Recomputed normals and depths equal
    the manifolds bit for bit; gradients of sum depth^2 + sum normal . w match
    the FP64 dual adjoint where FP64 EPA converges on the same face (normals
    within 1e-5) and FP64 central differences see one smooth branch."""
    sc = scenes.mixed(20000, seed=5, side=scenes.mixed_side_for(20000, 2.0))
    with physgpu.CollisionPipeline(sc.count, 8 * sc.count,
                                   max_hull_points=sc.hull_points.shape[1]) as pipe:
        pipe.set_colliders(sc)
        pipe.collide(algorithm=abi.UNIFORM_GRID)
        man = pipe.manifolds()
        normal, depth, point = pipe.contact_geometry()
        m = len(man)
        rng = np.random.default_rng(8)
        w = rng.standard_normal((3, m))
        w = (w / np.linalg.norm(w, axis=0)).astype(np.float32)
        used = np.arange(4)[:, None] < man["point_count"][None, :]
        gd = np.where(used, 2.0 * man["penetration"].T, 0.0).astype(np.float32)
        # Sphere-hull manifolds get zero gradients, reported by count and
        # by a warning.
        with pytest.warns(physgpu.AdjointSkippedWarning):
            got, skipped = pipe.contact_adjoint(w, gd, per_pair=True)
    cls = man["shape_pair"]
    poly = (cls == abi.PAIR_BOX_HULL) | (cls == abi.PAIR_HULL_HULL)
    assert poly.sum() > 500
    assert skipped == int((cls == abi.PAIR_SPHERE_HULL).sum()) > 0
    assert not got[:, cls == abi.PAIR_SPHERE_HULL].any()
    assert np.array_equal(normal.T[poly].view(np.uint32), man["normal"][poly].view(np.uint32))
    assert np.array_equal(depth[0][poly].view(np.uint32),
                          man["penetration"][poly, 0].view(np.uint32))
    pairs = _pairs_of(man)[poly]
    # FP64 runs stop at the LDS bounds (no continuation): the GPU adjoint
    # differentiates only runs inside them, and past them FP64 can settle on a
    # neighbouring near-coplanar face within the same-face normal tolerance.
    oracle.mirror_gpu_epa(continuation=False)
    try:
        want, status = oracle.contact_adjoint_f64(sc, pairs, w[:, poly], gd[:, poly], workers=16)
        fd, stable = oracle.contact_fd_f64(sc, pairs, w[:, poly], gd[:, poly], workers=16)
        n64, _, _, _ = oracle.contact_geometry_f64(sc, pairs)
    finally:
        oracle.mirror_gpu_epa(False)
    same_face = np.abs(n64.T - man["normal"][poly]).max(1) < 1e-5
    ok = (status == 1) & stable & same_face
    print(f"hull manifolds {poly.sum()}: compared {ok.sum()} (same FP64 face and smooth)")
    assert ok.mean() > 0.8
    rel = _rel(got[:, poly][:, ok].astype(np.float64), want[:, ok])
    print(f"GPU vs FP64 dual (EPA face): max rel {rel.max():.2e}")
    assert rel.max() <= TOL
    assert _rel(got[:, poly][:, ok].astype(np.float64), fd[:, ok]).max() <= TOL
