"""World-level collision section on the GPU with compound bodies: 1 to 4
colliders per body (physics_world.cpp:539-739).  Against the reference's own
PhysicsWorld::step (tests/golden/world_bodies.npz) for spheres and boxes, and
against the oracle with hulls; StepStats counters (physgpu_step_stats) against
both and against their closed forms."""
import numpy as np
import pytest

from conftest import assert_manifolds_equal, golden
from physgpu import abi, scenes

pytestmark = pytest.mark.gpu

ALGOS = [abi.SWEEP_AND_PRUNE, abi.UNIFORM_GRID, abi.DYNAMIC_TREE]


def _scene(g):
    return scenes.Scene(pose=g["pose"], shape_type=g["shape_type"], shape_data=g["shape_data"],
                        body_position=g["body_position"], local_pose=g["local_pose"],
                        friction=g["friction"], restitution=g["restitution"], body=g["body"])


def _possible(body, offsets=None):
    offsets = [0, len(body)] if offsets is None else offsets
    total = 0
    for w in range(len(offsets) - 1):
        b = body[offsets[w]:offsets[w + 1]]
        n = len(b)
        k = np.bincount(b) if n else np.zeros(0, np.int64)
        total += n * (n - 1) // 2 - int((k * (k - 1) // 2).sum())
    return total


@pytest.fixture(scope="module")
def pipe(physgpu):
    p = physgpu.CollisionPipeline(20000, 1 << 18, max_hull_points=16 * 20000)
    yield p
    p.close()


@pytest.fixture(scope="module", autouse=True)
def gpu_epa(oracle):
    oracle.mirror_gpu_epa()
    yield
    oracle.mirror_gpu_epa(False)


@pytest.mark.parametrize("algo", ALGOS)
def test_compound_bodies_match_reference(pipe, algo):
    g = golden("world_bodies.npz")
    sc = _scene(g)
    name = ("sap", "grid", "tree")[algo]
    print(f"compound bodies algo {algo}", flush=True)
    pipe.set_colliders(sc)
    pipe.set_world(sc)
    pipe.set_timing(True)
    pipe.collide(algorithm=algo)
    man = pipe.manifolds()
    st = pipe.step_stats()
    pipe.set_timing(False)
    got = man.copy()
    got["first"] = sc.body[man["first"]]  # the reference records body handles
    got["second"] = sc.body[man["second"]]
    assert_manifolds_equal(got, g["contacts_" + name], "compound bodies " + name)
    cand, possible, contacts = (int(v) for v in g["counts_" + name])
    assert st["candidate_pairs"] == cand
    assert st["possible_pairs"] == possible == _possible(sc.body)
    assert st["contact_count"] == contacts == len(man)
    nps = st["narrow_phase"]
    assert [nps["sphere_sphere_candidates"], nps["sphere_box_candidates"],
            nps["box_box_candidates"], nps["sphere_sphere_contacts"],
            nps["sphere_box_contacts"], nps["box_box_contacts"]] == \
        [int(v) for v in g["np_stats_" + name]]
    assert st["broad_phase"]["aabb_pairs"] >= cand
    assert st["broad_phase_ms"] > 0 and st["narrow_phase_ms"] > 0


def test_compound_bodies_with_hulls(pipe, oracle):
    """Hull colliders on compound bodies (the reference has no hull shape):
    GPU against the oracle bit for bit, body-local anchors included."""
    sc = scenes.compound_bodies(3000, seed=77, with_hulls=True)
    print(f"compound bodies with hulls: {sc.count} colliders", flush=True)
    want, t, _ = oracle.world_collide(sc, algorithm=abi.UNIFORM_GRID)
    pipe.set_colliders(sc)
    pipe.set_world(sc)
    pipe.collide(algorithm=abi.UNIFORM_GRID)
    man = pipe.manifolds()
    assert (man["shape_pair"] >= abi.PAIR_SPHERE_HULL).sum() > 100
    assert_manifolds_equal(man, want, "compound bodies with hulls")
    st = pipe.step_stats()
    assert st["candidate_pairs"] == t["candidate_pairs"]
    assert st["possible_pairs"] == int(t["possible_pairs"]) == _possible(sc.body)
    assert st["contact_count"] == len(want)


def test_possible_pairs_per_world(pipe):
    """Batched worlds: possible pairs are counted inside each world."""
    parts = [scenes.mixed(n, seed=90 + n, side=scenes.mixed_side_for(n, 1.0), with_hulls=False)
             for n in (17, 300, 1000, 64)]
    sc, off = scenes.concat(parts)
    pipe.set_colliders(sc)
    pipe.set_worlds(off)
    pipe.collide(algorithm=abi.UNIFORM_GRID)
    st = pipe.step_stats()
    pipe.set_worlds(None)
    assert st["possible_pairs"] == _possible(np.arange(sc.count), list(off))
    assert st["candidate_pairs"] == st["broad_phase"]["aabb_pairs"]
