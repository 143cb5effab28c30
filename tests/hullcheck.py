"""Brute-force FP64 check of GJK/EPA manifolds (pairs with a convex hull).

The hull contact is defined by this project (the reference has no hull shape,
include/phys/collision/shapes/convex_hull.hpp:5, and no caller of
computeEpa, src/collision/epa.cpp:64): one point at the midpoint of the two
deepest support points along the EPA normal n (from A toward B),

    w = 1/2 (s_A(n) + s_B(-n)),

with anchors the collider-local forms of w, and depth the EPA depth, which
at convergence equals the Minkowski support n . (s_A(n) - s_B(-n)) to within
the EPA tolerance.  This module recomputes every support point here, in
float64, by exhaustive search over each shape's features (hull points, the 8
box corners, sphere centre + r n/|n|), independently of the support maps of
the oracle and of the HIP kernels, and reports the manifolds that disagree.

Ties: a box (or a hull) can have several features at the same support value
(n parallel to a face or an edge); any tied feature is accepted.

Degenerate faces: where EPA converges on a face whose cross product is no
longer than 1e-8, the reference keeps that cross product unnormalised as the
face normal (epa.cpp:35-38), so the manifold's normal is a tiny vector and its
depth ~0 (touching contacts).  A sphere's support along such a direction is
its centre, because Math3d::normalize returns zero below 1e-8
(math_utils.cpp:19-23); the check follows both rules.
"""
import numpy as np

from physgpu import abi

CORNERS = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)],
                   np.float64)


def rotate(q, v):
    """Rotate v [..., 3] by the quaternion q (x, y, z, w), float64."""
    q = np.asarray(q, np.float64)
    u, w = q[:3], q[3]
    v = np.asarray(v, np.float64)
    return ((w * w - u @ u) * v + 2.0 * (v @ u)[..., None] * u + 2.0 * w * np.cross(u, v))


def features(scene, i):
    """Collider-local feature points of collider i ([k, 3] float64) and, for a
    sphere, its radius (features = centre only)."""
    t = int(scene.shape_type[i])
    if t == abi.SPHERE:
        return np.zeros((1, 3)), float(scene.shape_data[0, i])
    if t == abi.BOX:
        return CORNERS * scene.shape_data[:, i].astype(np.float64), 0.0
    f, c = scene.hull_range[:, i]
    return scene.hull_points[:, f:f + c].T.astype(np.float64), 0.0


def support_set(scene, i, d, tie=1e-5):
    """World support points of collider i along direction d: every feature
    within `tie` (m) of the maximum projection, float64 [k, 3]."""
    pose = scene.pose[:, i].astype(np.float64)
    local, radius = features(scene, i)
    off = rotate(pose[3:], local)            # offsets from the collider origin
    length = np.linalg.norm(d)
    proj = off @ (d / length)                # no 128 m cancellation here
    keep = proj >= proj.max() - tie
    pts = off[keep] + pose[:3]
    if radius and length > 1e-8:
        pts = pts + radius * (d / length)
    return pts


def check(scene, manifolds, epa_tolerance=1e-6, point_tol=2e-4, depth_tol=2e-4):
    """Check every hull manifold of `manifolds` (collider-local anchors, as
    physgpu_generate_contacts returns them).  Returns a dict of boolean arrays
    (one entry per hull manifold): anchors agree, point is the support
    midpoint, depth is the support distance; and the checked indices."""
    m = np.asarray(manifolds)
    hull = m["shape_pair"] >= abi.PAIR_SPHERE_HULL
    idx = np.nonzero(hull)[0]
    anchors = np.zeros(len(idx), bool)
    midpoint = np.zeros(len(idx), bool)
    depth_ok = np.zeros(len(idx), bool)
    for k, j in enumerate(idx):
        r = m[j]
        a, b = int(r["first"]), int(r["second"])
        assert r["point_count"] == 1
        pa = scene.pose[:, a].astype(np.float64)
        pb = scene.pose[:, b].astype(np.float64)
        wa = rotate(pa[3:], r["anchor_a"][0]) + pa[:3]
        wb = rotate(pb[3:], r["anchor_b"][0]) + pb[:3]
        anchors[k] = np.abs(wa - wb).max() <= point_tol
        n = r["normal"].astype(np.float64)
        sa = support_set(scene, a, n)
        sb = support_set(scene, b, -n)
        mids = 0.5 * (sa[:, None, :] + sb[None, :, :])
        midpoint[k] = np.abs(mids - wa).max(axis=2).min() <= point_tol
        # Minkowski support distance along n, the quantity EPA's convergence
        # test compares with the face distance (any tied pair gives the same).
        sd = n @ (sa[0] - sb[0])
        depth_ok[k] = abs(sd - float(r["penetration"][0])) <= epa_tolerance + depth_tol
    return dict(index=idx, anchors=anchors, midpoint=midpoint, depth=depth_ok)
