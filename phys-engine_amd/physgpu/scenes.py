"""Seeded synthetic collider scenes (numpy, float32, SoA planes).

Each scene holds exactly the arrays the C ABI takes (physgpu_colliders):
  pose        float32 [7, n]  px, py, pz, qx, qy, qz, qw (world pose)
  shape_type  uint8   [n]     0 sphere, 1 box, 2 hull
  shape_data  float32 [3, n]  sphere radius / box half extents
  hull_range  uint32  [2, n]  (first point, count) into hull_points, or None
  hull_points float32 [3, H]  collider-local hull points, or None
and, for world-level runs, body_position [3, n], local_pose [7, n],
friction [n], restitution [n].

Generators use numpy's PCG64 (np.random.default_rng), which gives identical
bytes for a given seed on every machine running the same numpy, so CPU and
GPU runs read the same inputs.  Workloads:
  spheres     4,096 spheres, r ~ U[0.25, 0.5], centres U[-10, 10]^3
  mixed       50% spheres r ~ U[0.2, 0.5], 40% boxes h ~ U[0.2, 0.6]^3,
              10% hulls of 16 points on an ellipsoid, semi-axes U[0.3, 0.6]^3,
              uniform unit quaternions, centres uniform in a cube of side L
  batched_worlds  worlds with body counts floor(exp(U[ln 16, ln 4096])), each a
              mixed scene in a cube of side L0 n^(1/3) (about one pair per
              body), concatenated with per-world collider offsets
  contact_pairs  isolated touching pairs, half box-box and half sphere-box /
              box-sphere, penetration depth U[1e-3, 5e-2] m, away from SAT
              axis ties and sphere-box region boundaries (adjoint checks)
  stack       64 x 64 columns of 64 unit boxes sunk 1e-3 m into each other,
              yaw U[-0.05, 0.05], lateral jitter U[-0.01, 0.01], on a ground
              box of half extents (100, 0.5, 100)
"""
from dataclasses import dataclass, field

import numpy as np

from . import abi


@dataclass
class Scene:
    pose: np.ndarray
    shape_type: np.ndarray
    shape_data: np.ndarray
    hull_range: np.ndarray = None
    hull_points: np.ndarray = None
    body_position: np.ndarray = None
    local_pose: np.ndarray = None
    friction: np.ndarray = None
    restitution: np.ndarray = None
    body: np.ndarray = None  # owning body per collider (world level); None: one each
    meta: dict = field(default_factory=dict)

    @property
    def count(self):
        return int(self.shape_type.shape[0])

    def subset(self, index):
        """Scene of the colliders `index` (hull points re-based)."""
        index = np.asarray(index, dtype=np.int64)
        hr = hp = None
        if self.hull_range is not None:
            hr = np.zeros((2, len(index)), np.uint32)
            pts = []
            cursor = 0
            for k, i in enumerate(index):
                if self.shape_type[i] == abi.HULL:
                    f, c = self.hull_range[:, i]
                    hr[:, k] = (cursor, c)
                    pts.append(self.hull_points[:, f:f + c])
                    cursor += int(c)
            hp = (np.concatenate(pts, axis=1) if pts else np.zeros((3, 1), np.float32))
            hp = np.ascontiguousarray(hp, dtype=np.float32)
        take = (lambda a: None if a is None else np.ascontiguousarray(a[..., index]))
        return Scene(pose=take(self.pose), shape_type=take(self.shape_type),
                     shape_data=take(self.shape_data), hull_range=hr, hull_points=hp,
                     body_position=take(self.body_position), local_pose=take(self.local_pose),
                     friction=take(self.friction), restitution=take(self.restitution),
                     body=take(self.body), meta=dict(self.meta, subset=len(index)))


def random_quaternions(rng, n):
    """Uniform unit quaternions (normalised 4-D Gaussians), float32 [4, n] xyzw."""
    q = rng.standard_normal((4, n))
    q /= np.linalg.norm(q, axis=0, keepdims=True)
    return q.astype(np.float32)


def _identity_quats(n):
    q = np.zeros((4, n), np.float32)
    q[3] = 1.0
    return q


def spheres(n=4096, seed=20261203, half_side=10.0, rmin=0.25, rmax=0.5):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-half_side, half_side, (3, n)).astype(np.float32)
    r = rng.uniform(rmin, rmax, n).astype(np.float32)
    pose = np.concatenate([pos, _identity_quats(n)]).astype(np.float32)
    data = np.zeros((3, n), np.float32)
    data[0] = r
    return Scene(pose=np.ascontiguousarray(pose), shape_type=np.zeros(n, np.uint8),
                 shape_data=data, meta=dict(workload="spheres", n=n, seed=seed))


def mixed_side_for(n, pairs_per_body=1.0):
    """Cube side giving about `pairs_per_body` candidate pairs per body for the
    mixed shape mix (calibrated: 1,048,576 bodies in a 256 m cube give ~0.26
    pairs per body; pair density scales with n / L^3)."""
    base_pairs_per_body = 0.263
    side = 256.0 * (n / 1048576.0) ** (1.0 / 3.0)
    return side * (base_pairs_per_body / pairs_per_body) ** (1.0 / 3.0)


def mixed(n=1048576, seed=20261203, side=256.0, hull_points=16, with_hulls=True):
    rng = np.random.default_rng(seed)
    n_sph = n // 2
    n_box = int(round(0.4 * n)) if with_hulls else n - n_sph
    n_hull = n - n_sph - n_box
    types = np.concatenate([np.zeros(n_sph, np.uint8), np.ones(n_box, np.uint8),
                            np.full(n_hull, abi.HULL, np.uint8)])
    types = types[rng.permutation(n)]
    pos = rng.uniform(-side / 2, side / 2, (3, n)).astype(np.float32)
    quat = random_quaternions(rng, n)
    data = np.zeros((3, n), np.float32)
    sph = types == abi.SPHERE
    box = types == abi.BOX
    hull = types == abi.HULL
    data[0, sph] = rng.uniform(0.2, 0.5, sph.sum())
    data[:, box] = rng.uniform(0.2, 0.6, (3, box.sum()))
    hr = hp = None
    if n_hull:
        idx = np.nonzero(hull)[0]
        axes = rng.uniform(0.3, 0.6, (3, n_hull))
        dirs = rng.standard_normal((n_hull, hull_points, 3))
        dirs /= np.linalg.norm(dirs, axis=2, keepdims=True)
        pts = dirs * axes.T[:, None, :]
        hp = np.ascontiguousarray(pts.reshape(-1, 3).T.astype(np.float32))
        hr = np.zeros((2, n), np.uint32)
        hr[0, idx] = np.arange(n_hull, dtype=np.uint32) * hull_points
        hr[1, idx] = hull_points
    pose = np.ascontiguousarray(np.concatenate([pos, quat]).astype(np.float32))
    return Scene(pose=pose, shape_type=types, shape_data=data, hull_range=hr, hull_points=hp,
                 meta=dict(workload="mixed", n=n, seed=seed, side=side))


def concat(parts):
    """One scene holding `parts` back to back (hull ranges re-based) and the
    collider offsets of each part (len(parts) + 1 entries)."""
    counts = np.array([p.count for p in parts], np.int64)
    offsets = np.zeros(len(parts) + 1, np.uint32)
    offsets[1:] = np.cumsum(counts)
    n = int(offsets[-1])
    pose = np.concatenate([p.pose for p in parts], axis=1)
    types = np.concatenate([p.shape_type for p in parts])
    data = np.concatenate([p.shape_data for p in parts], axis=1)
    hr = hp = None
    if any(p.hull_range is not None for p in parts):
        hr = np.zeros((2, n), np.uint32)
        pts = []
        cursor = 0
        for p, o in zip(parts, offsets[:-1]):
            if p.hull_range is None:
                continue
            hr[:, o:o + p.count] = p.hull_range
            hr[0, o:o + p.count] += np.uint32(cursor)
            pts.append(p.hull_points)
            cursor += p.hull_points.shape[1]
        hp = np.ascontiguousarray(np.concatenate(pts, axis=1), np.float32)
    sc = Scene(pose=np.ascontiguousarray(pose), shape_type=np.ascontiguousarray(types),
               shape_data=np.ascontiguousarray(data), hull_range=hr, hull_points=hp,
               meta=dict(workload="concat", n=n, parts=len(parts)))
    return sc, offsets


def batched_worlds(worlds=4096, seed=20261203, min_bodies=16, max_bodies=4096,
                   pairs_per_body=1.33):
    """Independent mixed worlds of uneven size, concatenated.

    World w has n_w = floor(exp(U[ln min, ln max])) bodies drawn like `mixed`
    (seeded by (seed, w)) in a cube of side mixed_side_for(n_w, pairs_per_body)
    = L0 n_w^(1/3).  `pairs_per_body` is the bulk density; the cube walls cut
    the pairs of small worlds, and the default 1.33 gives about 1.0 candidate
    pair (2 overlaps) per body over the batch.  Returns (scene, offsets, counts)."""
    rng = np.random.default_rng(seed)
    counts = np.floor(np.exp(rng.uniform(np.log(min_bodies), np.log(max_bodies),
                                         worlds))).astype(np.int64)
    parts = [mixed(int(c), seed=[seed, w], side=mixed_side_for(int(c), pairs_per_body))
             for w, c in enumerate(counts)]
    sc, offsets = concat(parts)
    sc.meta = dict(workload="batched_worlds", worlds=worlds, n=sc.count, seed=seed,
                   min_bodies=min_bodies, max_bodies=max_bodies, pairs_per_body=pairs_per_body)
    return sc, offsets, counts


def _rotate(q, v):
    """Rotate v [3, n] by unit quaternions q [4, n] (xyzw), float64."""
    u, w = q[:3], q[3]
    return ((w * w - (u * u).sum(0)) * v + 2.0 * (u * v).sum(0) * u +
            2.0 * w * np.cross(u, v, axis=0))


def _box_axes(q):
    """Columns of the rotation matrix of q [4, n]: [3 (column), 3, n]."""
    eye = np.eye(3)
    return np.stack([_rotate(q, np.repeat(eye[:, k:k + 1], q.shape[1], 1)) for k in range(3)])


def contact_pairs(pairs=65536, seed=20261203, spacing=8.0, tie_margin=1e-3, inside_fraction=0.0):
    """Isolated touching pairs for adjoint checks: colliders (2i, 2i+1).

    Even pairs are box-box: random orientations, B offset from A along a random
    direction by the distance that gives SAT depth U[1e-3, 5e-2]; pairs whose
    best and second-best SAT axes differ in depth by less than
    tie_margin x centre distance (a 1e-3 rad turn cannot swap them) are drawn
    again.  Odd pairs are sphere-box (sphere first on pairs 1 mod 4, box first
    on 3 mod 4): the sphere centre sits outside the box at distance r - depth
    from its closest point, at least 1e-3 m from the boundaries of the box's
    face / edge / corner regions; with inside_fraction > 0 that share of the
    sphere-box pairs instead has the sphere centre inside the box (depth
    r + distance to the nearest face, nearest and second-nearest faces at
    least 1e-3 m apart).  Pairs sit on a lattice `spacing` apart."""
    rng = np.random.default_rng(seed)
    n = 2 * pairs
    side = int(np.ceil(pairs ** (1.0 / 3.0)))
    g = np.stack(np.meshgrid(np.arange(side), np.arange(side), np.arange(side),
                             indexing="ij")).reshape(3, -1)[:, :pairs].astype(np.float64)
    centre = (g - (side - 1) / 2.0) * spacing
    bb = np.arange(pairs) % 2 == 0
    pose = np.zeros((7, n), np.float64)
    types = np.zeros(n, np.uint8)
    data = np.zeros((3, n), np.float64)

    # Box-box.
    idx = np.nonzero(bb)[0]
    todo = idx.copy()
    while len(todo):
        k = len(todo)
        qa = random_quaternions(rng, k).astype(np.float64)
        qb = random_quaternions(rng, k).astype(np.float64)
        ha = rng.uniform(0.2, 0.6, (3, k))
        hb = rng.uniform(0.2, 0.6, (3, k))
        u = rng.standard_normal((3, k))
        u /= np.linalg.norm(u, axis=0)
        depth = rng.uniform(1e-3, 5e-2, k)
        ea, eb = _box_axes(qa), _box_axes(qb)
        axes = [ea[i] for i in range(3)] + [eb[i] for i in range(3)]
        axes += [np.cross(ea[i], eb[j], axis=0) for i in range(3) for j in range(3)]
        s_k, d_fn = [], []
        for a in axes:
            ln = np.linalg.norm(a, axis=0)
            ok = ln ** 2 >= 1e-8
            a = a / np.where(ok, ln, 1.0)
            ra = sum(np.abs((ea[c] * a).sum(0)) * ha[c] for c in range(3))
            rb = sum(np.abs((eb[c] * a).sum(0)) * hb[c] for c in range(3))
            ua = np.abs((u * a).sum(0))
            s = np.where(ok & (ua > 1e-12), (ra + rb - depth) / np.maximum(ua, 1e-12), np.inf)
            s_k.append(s)
            d_fn.append((np.where(ok, ra + rb, np.inf), ua))
        S = np.stack(s_k)
        s_star = S.min(0)
        depths = np.stack([r - s_star * ua for r, ua in d_fn])
        srt = np.sort(depths, axis=0)
        good = (np.isfinite(s_star) & (s_star > 0) &
                (srt[1] - srt[0] >= tie_margin * s_star) & (np.abs(srt[0] - depth) < 1e-9))
        ids = todo[good]
        a_i, b_i = 2 * ids, 2 * ids + 1
        pose[:3, a_i] = centre[:, ids]
        pose[3:, a_i] = qa[:, good]
        pose[:3, b_i] = centre[:, ids] + s_star[good] * u[:, good]
        pose[3:, b_i] = qb[:, good]
        types[a_i] = types[b_i] = 1
        data[:, a_i] = ha[:, good]
        data[:, b_i] = hb[:, good]
        todo = todo[~good]

    # Sphere-box.
    idx = np.nonzero(~bb)[0]
    todo = idx.copy()
    while len(todo):
        k = len(todo)
        q = random_quaternions(rng, k).astype(np.float64)
        h = rng.uniform(0.2, 0.6, (3, k))
        r = rng.uniform(0.2, 0.5, k)
        depth = rng.uniform(1e-3, 5e-2, k)
        x = rng.uniform(-1.0, 1.0, (3, k)) * (h + r)
        closest = np.clip(x, -h, h)
        d = x - closest
        dl = np.linalg.norm(d, axis=0)
        local = closest + d / np.where(dl > 0, dl, 1.0) * (r - depth)
        margin = np.abs(np.abs(local) - h).min(0)
        good = (dl > 1e-3) & (margin >= 1e-3)
        inside = rng.uniform(0.0, 1.0, k) < inside_fraction
        if inside.any():
            xin = rng.uniform(-0.9, 0.9, (3, k)) * h
            faces = np.sort(np.concatenate([h - xin, h + xin]), axis=0)
            local = np.where(inside, xin, local)
            good = np.where(inside, faces[1] - faces[0] >= 1e-3, good)
        ids = todo[good]
        sphere_first = (ids % 4) == 1
        box_i = np.where(sphere_first, 2 * ids + 1, 2 * ids)
        sph_i = np.where(sphere_first, 2 * ids, 2 * ids + 1)
        pose[:3, box_i] = centre[:, ids]
        pose[3:, box_i] = q[:, good]
        types[box_i] = 1
        data[:, box_i] = h[:, good]
        pose[:3, sph_i] = centre[:, ids] + _rotate(q[:, good], local[:, good])
        pose[6, sph_i] = 1.0
        types[sph_i] = 0
        data[0, sph_i] = r[good]
        todo = todo[~good]
    return Scene(pose=np.ascontiguousarray(pose.astype(np.float32)), shape_type=types,
                 shape_data=np.ascontiguousarray(data.astype(np.float32)),
                 meta=dict(workload="contact_pairs", pairs=pairs, n=n, seed=seed))


SEGMENT_ROWS = (2, 4, 5, 16, 17, 63, 64, 65, 100, 400)
SEGMENT_SLABS = ((63, 65), (64, 64), (41, 100), (100, 250))


def pair_segments(nx=400, ny=250, nz=3, seed=20261203, rows=SEGMENT_ROWS, slabs=SEGMENT_SLABS):
    """Pairs per body of chosen sizes: spheres r = 0.4 on the unit lattice
    nx x ny x nz (no two overlap) in shuffled order after 1 + len(rows) +
    len(slabs) boxes: body 0 covers the whole lattice, row box k covers
    rows[k] lattice points along x in layer 0, slab box k covers a
    slabs[k][0] x slabs[k][1] patch (layer 1, the last one layer 2).  Boxes
    overlap nothing but spheres and box 0, so box k's pairs as `first` are
    exactly the points it covers and box 0 has one pair with every other body.
    meta["segments"] lists the expected count per box."""
    rng = np.random.default_rng(seed)
    boxes = []  # (centre, half extents, expected pairs)
    for k, length in enumerate(rows):
        y = 2 * k
        assert y < ny and length <= nx
        boxes.append(((length - 1) / 2.0, y, 0.0, (length - 1) / 2.0 + 0.45, 0.45, 0.45, length))
    x0 = 0
    for k, (a, b) in enumerate(slabs):
        layer = 1 if k + 1 < len(slabs) else 2
        if layer == 2:
            x0 = 0
        assert x0 + a <= nx and b <= ny and layer < nz
        boxes.append((x0 + (a - 1) / 2.0, (b - 1) / 2.0, float(layer),
                      (a - 1) / 2.0 + 0.45, (b - 1) / 2.0 + 0.45, 0.45, a * b))
        x0 += a + 1
    lattice = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"),
                       0).reshape(3, -1).astype(np.float32)
    lattice = lattice[:, rng.permutation(lattice.shape[1])]
    m = lattice.shape[1]
    nb = 1 + len(boxes)
    n = nb + m
    pose = np.zeros((7, n), np.float32)
    pose[6] = 1.0
    data = np.zeros((3, n), np.float32)
    types = np.zeros(n, np.uint8)
    types[:nb] = abi.BOX
    pose[:3, 0] = ((nx - 1) / 2.0, (ny - 1) / 2.0, (nz - 1) / 2.0)
    data[:, 0] = (nx / 2.0, ny / 2.0, nz / 2.0)
    for k, (cx, cy, cz, hx, hy, hz, _) in enumerate(boxes):
        pose[:3, 1 + k] = (cx, cy, cz)
        data[:, 1 + k] = (hx, hy, hz)
    pose[:3, nb:] = lattice
    data[0, nb:] = 0.4
    segments = [n - 1] + [b[-1] for b in boxes]
    return Scene(pose=pose, shape_type=types, shape_data=data,
                 meta=dict(workload="pair_segments", n=n, seed=seed, segments=segments))


def stack(columns=64, height=64, seed=20261203, spacing=1.1, sink=1e-3):
    """Resting stack: box k of a column at y = 0.5 + k (1 - sink) over a ground box."""
    rng = np.random.default_rng(seed)
    n = columns * columns * height + 1
    cx, cz, k = np.meshgrid(np.arange(columns), np.arange(columns), np.arange(height),
                            indexing="ij")
    cx, cz, k = cx.ravel(), cz.ravel(), k.ravel()
    m = cx.size
    off = (columns - 1) * spacing / 2.0
    x = cx * spacing - off + rng.uniform(-0.01, 0.01, m)
    z = cz * spacing - off + rng.uniform(-0.01, 0.01, m)
    y = 0.5 + k * (1.0 - sink)
    yaw = rng.uniform(-0.05, 0.05, m)
    pose = np.zeros((7, n), np.float32)
    pose[0, :m], pose[1, :m], pose[2, :m] = x, y, z
    pose[4, :m] = np.sin(yaw / 2)
    pose[6, :m] = np.cos(yaw / 2)
    pose[:, m] = (0, -0.5, 0, 0, 0, 0, 1)
    data = np.full((3, n), 0.5, np.float32)
    data[:, m] = (100.0, 0.5, 100.0)
    return Scene(pose=pose, shape_type=np.ones(n, np.uint8), shape_data=data,
                 meta=dict(workload="stack", n=n, seed=seed, columns=columns, height=height))


def with_world_extras(scene, seed=7, local_rotation=True):
    """World-level form of a scene: each collider on its own body at the pose
    position with identity body rotation, the collider's orientation carried by
    its local transform (the reference has no body-rotation setter), and
    per-body friction/restitution."""
    rng = np.random.default_rng(seed)
    n = scene.count
    scene.body_position = np.ascontiguousarray(scene.pose[:3].copy())
    lp = np.zeros((7, n), np.float32)
    lp[6] = 1.0
    if local_rotation:
        lp[3:7] = scene.pose[3:7]
    scene.local_pose = lp
    scene.friction = rng.uniform(-0.1, 1.0, n).astype(np.float32)
    scene.restitution = rng.uniform(0.0, 0.8, n).astype(np.float32)
    return scene


def compound_bodies(bodies=2000, seed=20261207, with_hulls=True, pairs_per_body=1.5):
    """World-level scene of compound bodies: body b carries 1 to 4 colliders
    (the mixed shape mix) at local offsets U[-0.6, 0.6]^3 with random local
    orientations; bodies sit at identity rotation (the reference has no
    body-rotation setter) in a cube sized for about `pairs_per_body` candidate
    pairs per collider, with per-body friction U[-0.1, 1] and restitution
    U[0, 0.8].  Colliders are listed body by body, so body ids appear in
    increasing order (the reference's body handle order)."""
    rng = np.random.default_rng(seed)
    per = rng.integers(1, 5, bodies)
    n = int(per.sum())
    body = np.repeat(np.arange(bodies, dtype=np.uint32), per)
    shapes = mixed(n, seed=seed + 1, side=1.0, with_hulls=with_hulls)
    side = mixed_side_for(n, pairs_per_body)
    bpos = rng.uniform(-side / 2, side / 2, (3, bodies)).astype(np.float32)
    lp = np.zeros((7, n), np.float32)
    lp[:3] = rng.uniform(-0.6, 0.6, (3, n))
    lp[3:7] = random_quaternions(rng, n)
    sc = Scene(pose=np.zeros((7, n), np.float32), shape_type=shapes.shape_type,
               shape_data=shapes.shape_data, hull_range=shapes.hull_range,
               hull_points=shapes.hull_points,
               meta=dict(workload="compound_bodies", bodies=bodies, n=n, seed=seed))
    sc.body = body
    sc.body_position = np.ascontiguousarray(bpos[:, body])
    sc.local_pose = lp
    # Collider world pose = body pose (identity rotation) composed with the
    # local transform: position body + local (the rotation is the local one).
    sc.pose = np.ascontiguousarray(np.concatenate([sc.body_position + lp[:3], lp[3:7]]),
                                   np.float32)
    fr = rng.uniform(-0.1, 1.0, bodies).astype(np.float32)
    rs = rng.uniform(0.0, 0.8, bodies).astype(np.float32)
    sc.friction = np.ascontiguousarray(fr[body])
    sc.restitution = np.ascontiguousarray(rs[body])
    return sc
