"""Write the golden fixtures of this directory from the reference engine.

Runs only where the reference sources exist (/root/reference): it builds
oracle/_ref/libphysref.so (the reference compiled from its own sources, see
oracle/Makefile) and records, for seeded inputs, what the reference itself
returns.  The fixtures hold inputs and expected outputs only.

  spheres.npz   4,096 spheres r ~ U[0.25, 0.5] in [-10, 10]^3: AABBs, SAP /
                grid / tree pair sets and stats, manifolds of every pair,
                world-level contacts
  mixed.npz     4,096 spheres and boxes with random orientations
  stack.npz     8 x 8 columns of 8 boxes on a ground box (face-face, 4 points)
  kat.npz       the narrow-phase known-answer cases of
                tests/collision/narrowphase_tests.cpp:35-179
  bounds.npz    boundary AABB sets of tests/collision/broadphase_tests.cpp:89-182
                and 30 seeded 192-box scenes x 4 motion steps (:107-138 style)
  gjk.npz       GJK + EPA (status, normal, depth, final simplex) over hull,
                box and sphere pairs with this project's support maps
  world.npz     world-level collision section (body-local anchors, materials)
  world_bodies.npz  world-level collision section of compound bodies (1 to 4
                colliders per body, spheres and boxes): same-body filter,
                StepStats candidatePairs / possiblePairs, per algorithm
  large.json    pair counts and SHA-256 of the pair lists at 1,048,576 mixed
                bodies and the 262,145-box stack
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "phys-engine_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from physgpu import abi, scenes  # noqa: E402


def scene_arrays(sc, prefix=""):
    out = {prefix + "pose": sc.pose, prefix + "shape_type": sc.shape_type,
           prefix + "shape_data": sc.shape_data}
    if sc.hull_range is not None:
        out[prefix + "hull_range"] = sc.hull_range
        out[prefix + "hull_points"] = sc.hull_points
    for k in ("body_position", "local_pose", "friction", "restitution"):
        v = getattr(sc, k)
        if v is not None:
            out[prefix + k] = v
    return out


def broad_and_narrow(ref, sc, with_world=False):
    out = scene_arrays(sc)
    aabbs = ref.compute_aabbs(sc)
    out["aabbs"] = aabbs
    for algo, name in ((0, "sap"), (1, "grid"), (2, "tree")):
        pairs, st = ref.find_pairs(aabbs, algo)
        out[f"pairs_{name}"] = pairs
        out[f"stats_{name}"] = np.array([st[k] for k, _ in abi.BroadPhaseStats._fields_],
                                        np.float64)
    pairs = out["pairs_sap"]
    rec, hit = ref.generate_contacts(sc, pairs)
    out["contacts"] = rec
    out["hit"] = hit
    if with_world:
        m, t, s = ref.world_collide(scenes.with_world_extras(sc))
        out.update(scene_arrays(sc, "w_"))
        out["world_contacts"] = m
        out["world_np_stats"] = s
        out["world_candidates"] = np.array([t["candidate_pairs"]], np.int64)
    return out


def kat_cases():
    """Inputs of narrowphase_tests.cpp:35-179 as (pose A, shape A, pose B, shape B)."""
    def q_axis(axis, angle):
        axis = np.asarray(axis, np.float32)
        half = np.float32(angle) * np.float32(0.5)
        s = np.float32(np.sin(half))
        return [axis[0] * s, axis[1] * s, axis[2] * s, np.float32(np.cos(half))]

    ident = [0, 0, 0, 1]
    cases = []
    # sphere-sphere r=1, r=1 at distance 1.5
    cases.append(([0, 0, 0] + ident, (0, [1, 0, 0]), [1.5, 0, 0] + ident, (0, [1, 0, 0])))
    # box rotated pi/2 about z vs sphere at (0, 1.5, 0)
    cases.append(([0, 0, 0] + q_axis([0, 0, 1], 1.57079632679), (1, [1, 1, 1]),
                  [0, 1.5, 0] + ident, (0, [1, 0, 0])))
    # box-box normal unit length
    cases.append(([0, 0, 0] + q_axis([0, 1, 0], 0.4), (1, [1, 1, 1]),
                  [1.2, 0.8, 0.6] + q_axis([1, 0, 0], 0.7), (1, [1, 1, 1])))
    # tilted box on a floor, both orders
    for angle, height in ((0.3, 0.5 * (np.sin(np.float32(0.3)) + np.cos(np.float32(0.3))) - 0.02),
                          (0.1, 0.3)):
        floor = ([0, -0.5, 0] + q_axis([0, 0, 1], 0.0), (1, [9, 0.5, 4]))
        box = ([0, np.float32(height), 0] + q_axis([0, 0, 1], angle), (1, [0.5, 0.5, 0.5]))
        cases.append(floor + box)
        cases.append(box + floor)
    return cases


def kat_scene():
    cases = kat_cases()
    n = 2 * len(cases)
    pose = np.zeros((7, n), np.float32)
    types = np.zeros(n, np.uint8)
    data = np.zeros((3, n), np.float32)
    for k, (pa, sa, pb, sb) in enumerate(cases):
        for j, (p, s) in enumerate(((pa, sa), (pb, sb))):
            i = 2 * k + j
            pose[:, i] = np.asarray(p, np.float32)
            types[i] = s[0]
            data[:, i] = np.asarray(s[1], np.float32)
    pairs = np.array([[2 * k, 2 * k + 1] for k in range(len(cases))], np.uint32)
    return scenes.Scene(pose=pose, shape_type=types, shape_data=data), pairs


def bounds_sets():
    F = np.float32
    big = np.finfo(np.float32).max
    sets = {
        "empty": np.zeros((0, 6), F),
        "point": np.zeros((1, 6), F),
        "identical16": np.tile(np.array([-1, -1, -1, 1, 1, 1], F), (16, 1)),
        "touching": np.array([[4, 0, 0, 5, 1, 1], [-2, -2, -2, 0, 0, 0], [0, 0, 0, 1, 1, 1],
                              [1, 0, 0, 2, 1, 1], [1, 1, 0, 2, 2, 1], [1, 1, 1, 2, 2, 2],
                              [0, 4, 0, 1, 5, 1], [0, 0, 4, 1, 1, 5], [1, 1, 1, 1, 1, 1],
                              [-100, -0.5, -100, 100, 0, 100]], F),
        "grid_two": np.tile(np.array([-0.5, -0.5, -0.5, 0.5, 0.5, 0.5], F), (2, 1)),
        "grid_limit": np.array([[-0.5, -0.5, -0.5, 0.5, 0.5, 0.5]] * 3 +
                               [[0, 0, 0, 3, 3, 3], [0, 0, 0, 4, 3, 3]], F),
        "extremes": np.array([[-0.5, -0.5, -0.5, 0.5, 0.5, 0.5],
                              [0.5, -0.5, -0.5, 1.5, 0.5, 0.5],
                              [-1.5, -0.5, -0.5, -0.5, 0.5, 0.5],
                              [-100000, -0.5, -100000, 100000, 0, 100000],
                              [big, 0, 0, big, 0, 0], [big, 0, 0, big, 0, 0],
                              [-2147483648.0, 0, 0, -2147483648.0, 0, 0],
                              [-2147483648.0, 0, 0, -2147483648.0, 0, 0]], F),
        "huge": np.array([[-1, -1, -1, 1, 1, 1], [-big, -big, -big, big, big, big]], F),
    }
    rng = np.random.default_rng(42)
    for scene in range(30):
        centre = rng.uniform(-20, 20, (192, 3)).astype(F)
        if scene % 3 == 0:
            centre = (centre * F(0.1)).astype(F)
        half = rng.uniform(0.01, 3.0, (192, 3)).astype(F)
        box = np.concatenate([centre - half, centre + half], axis=1).astype(F)
        if scene % 3 == 1:
            box[:, 0] = -1.0
            box[:, 3] = 1.0
        for step in range(4):
            sets[f"random{scene:02d}_{step}"] = box.copy()
            motion = (rng.uniform(-20, 20, (192, 3)) * 0.05).astype(F)
            box[:, :3] += motion
            box[:, 3:] += motion
    return sets


def pair_hash(pairs):
    return hashlib.sha256(np.ascontiguousarray(pairs, np.uint32).tobytes()).hexdigest()


def world_bodies(ref):
    """Compound bodies through the reference's PhysicsWorld::step."""
    sc = scenes.compound_bodies(2000, seed=20261207, with_hulls=False)
    out = scene_arrays(sc)
    out["body"] = sc.body
    for algo, name in ((0, "sap"), (1, "grid"), (2, "tree")):
        m, t, st = ref.world_collide(sc, algorithm=algo)
        out[f"contacts_{name}"] = m
        out[f"np_stats_{name}"] = st
        out[f"counts_{name}"] = np.array([t["candidate_pairs"], t["possible_pairs"], len(m)],
                                         np.int64)
    return out


def main(large=True, only=None):
    O.build(ref=True)
    ref = O.CpuCollision("ref")
    save = lambda name, d: np.savez_compressed(os.path.join(HERE, name), **d)  # noqa: E731
    if only == "world_bodies":
        save("world_bodies.npz", world_bodies(ref))
        print("world_bodies.npz written")
        return
    orc = O.CpuCollision("oracle")
    save("world_bodies.npz", world_bodies(ref))

    save("spheres.npz", broad_and_narrow(ref, scenes.spheres(4096, seed=20261203),
                                         with_world=True))
    mixed = scenes.mixed(4096, seed=20261204, side=scenes.mixed_side_for(4096, 1.0),
                         with_hulls=False)
    save("mixed.npz", broad_and_narrow(ref, mixed, with_world=True))
    save("stack.npz", broad_and_narrow(ref, scenes.stack(columns=8, height=8, seed=20261205)))

    sc, pairs = kat_scene()
    rec, hit = ref.generate_contacts(sc, pairs)
    save("kat.npz", dict(scene_arrays(sc), pairs=pairs, contacts=rec, hit=hit))

    bs = {}
    for name, b in bounds_sets().items():
        bs[f"{name}__bounds"] = b
        for algo, an in ((0, "sap"), (1, "grid"), (2, "tree")):
            p, st = ref.find_pairs(b, algo)
            if algo == 0:
                bs[f"{name}__pairs"] = p
            else:  # the three reference methods return one pair set
                assert np.array_equal(p, bs[f"{name}__pairs"])
            bs[f"{name}__{an}_stats"] = np.array(
                [st[k] for k, _ in abi.BroadPhaseStats._fields_], np.float64)
    save("bounds.npz", bs)

    hs = scenes.mixed(2048, seed=20261206, side=scenes.mixed_side_for(2048, 1.0))
    aabbs = orc.compute_aabbs(hs)  # hull AABBs are defined by this project
    hp, _ = ref.find_pairs(aabbs, 0)
    st, nrm, dep, smp = ref.gjk_epa(hs, hp, 64, 64, 1e-6)
    st2, nrm2, dep2, smp2 = ref.gjk_epa(hs, hp, 32, 64, 1e-4)  # the reference defaults
    save("gjk.npz", dict(scene_arrays(hs), pairs=hp, status=st, normal=nrm, depth=dep,
                         simplex=smp, d_status=st2, d_normal=nrm2, d_depth=dep2,
                         d_simplex=smp2))

    if large:
        info = {}
        big = scenes.mixed(1048576, seed=20261203, side=256.0)
        ab = orc.compute_aabbs(big)
        nh = big.shape_type != abi.HULL
        rab = ref.compute_aabbs(big.subset(np.nonzero(nh)[0]))
        assert np.array_equal(rab.view(np.uint32), ab[nh].view(np.uint32))
        p, st = ref.find_pairs(ab, 1)
        info["mixed1M"] = dict(n=big.count, seed=20261203, side=256.0, pairs=int(len(p)),
                               sha256=pair_hash(p), grid_cell=st["grid_cell_size"])
        stk = scenes.stack(columns=64, height=64, seed=20261203)
        ab = ref.compute_aabbs(stk)
        p, st = ref.find_pairs(ab, 1)
        rec, hit = ref.generate_contacts(stk, p)
        info["stack262145"] = dict(n=stk.count, seed=20261203, pairs=int(len(p)),
                                   sha256=pair_hash(p), manifolds=int(hit.sum()),
                                   points=int(rec["point_count"][hit].sum()))
        with open(os.path.join(HERE, "large.json"), "w") as f:
            json.dump(info, f, indent=1)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    only = sys.argv[sys.argv.index("--only") + 1] if "--only" in sys.argv else None
    main(large="--no-large" not in sys.argv, only=only)
