"""HIP path against the CPU oracle, bit for bit, on seeded scenes small enough
for the oracle to finish in seconds.  Every call goes through the C ABI
(libphysgpu.so) via the physgpu host module."""
import numpy as np
import pytest

from conftest import assert_manifolds_equal, compact
from physgpu import abi, scenes

pytestmark = pytest.mark.gpu

ALGOS = [abi.SWEEP_AND_PRUNE, abi.UNIFORM_GRID, abi.DYNAMIC_TREE]


def _cases():
    return {
        "spheres4096": scenes.spheres(4096, seed=11),
        "mixed4096": scenes.mixed(4096, seed=12, side=scenes.mixed_side_for(4096, 1.0),
                                  with_hulls=False),
        "hulls4096": scenes.mixed(4096, seed=13, side=scenes.mixed_side_for(4096, 1.0)),
        "stack8": scenes.stack(columns=8, height=8, seed=14),
    }


CASES = _cases()


@pytest.fixture(scope="module", autouse=True)
def bounded_oracle(oracle):
    """Expected results use the GPU's EPA polytope bounds."""
    oracle.mirror_gpu_epa()
    yield
    oracle.mirror_gpu_epa(False)


@pytest.fixture(scope="module")
def pipe(physgpu):
    p = physgpu.CollisionPipeline(70000, 1 << 20, max_hull_points=16 * 70000)
    yield p
    p.close()


@pytest.mark.parametrize("name", list(CASES))
def test_aabbs_bitwise(pipe, oracle, name):
    sc = CASES[name]
    print(f"case {name}", flush=True)
    pipe.set_colliders(sc)
    got = pipe.compute_aabbs()
    want = oracle.compute_aabbs(sc)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("algo", ALGOS)
@pytest.mark.parametrize("name", list(CASES))
def test_pairs_exact(pipe, oracle, name, algo):
    sc = CASES[name]
    print(f"case {name} algo {algo}", flush=True)
    pipe.set_colliders(sc)
    aabbs = pipe.compute_aabbs()
    got, stats = pipe.find_candidate_pairs(algorithm=algo, with_stats=True)
    want, ostats = oracle.find_pairs(aabbs, algo)
    assert got.shape == want.shape
    assert np.array_equal(got, want)
    assert stats["aabb_pairs"] == len(want)
    if algo == abi.SWEEP_AND_PRUNE:
        assert stats["x_window_comparisons"] == ostats["x_window_comparisons"]


@pytest.mark.parametrize("name", list(CASES))
def test_contacts_bitwise(pipe, oracle, name):
    sc = CASES[name]
    print(f"case {name}", flush=True)
    pipe.set_colliders(sc)
    aabbs = oracle.compute_aabbs(sc)
    pairs, _ = oracle.find_pairs(aabbs)
    rec, hit = oracle.generate_contacts(sc, pairs)
    got, stats = pipe.generate_contacts(pairs, with_stats=True)
    assert_manifolds_equal(got, compact(rec, hit), name)
    assert np.array_equal(got["pair_index"], np.nonzero(hit)[0].astype(np.uint32))
    st, _, _, _ = oracle.gjk_epa(sc, pairs)
    hull = (sc.shape_type[pairs[:, 0]] == abi.HULL) | (sc.shape_type[pairs[:, 1]] == abi.HULL)
    assert stats["epa_unresolved"] == int(((st == 3) & hull).sum())
    assert stats["gjk_epa_failures"] == int(((st == 2) & hull).sum())
    oracle.mirror_gpu_epa(continuation=False)
    try:
        sb, _, _, _ = oracle.gjk_epa(sc, pairs)
    finally:
        oracle.mirror_gpu_epa()
    assert stats["epa_overflows"] == int(((sb == 3) & hull).sum())
    for cls, key in ((0, "sphere_sphere"), (1, "sphere_box"), (2, "box_box")):
        assert stats[key + "_contacts"] == int((hit & (rec["shape_pair"] == cls)).sum())


@pytest.mark.parametrize("algo", ALGOS)
@pytest.mark.parametrize("name", list(CASES))
def test_collide_pipeline(pipe, oracle, name, algo):
    sc = CASES[name]
    print(f"case {name} algo {algo}", flush=True)
    pipe.set_colliders(sc)
    npairs, nman = pipe.collide(algorithm=algo)
    pairs = pipe.pairs()
    man = pipe.manifolds()
    aabbs = oracle.compute_aabbs(sc)
    want_pairs, _ = oracle.find_pairs(aabbs, algo)
    assert np.array_equal(pairs, want_pairs)
    rec, hit = oracle.generate_contacts(sc, want_pairs)
    assert_manifolds_equal(man, compact(rec, hit), name)
    assert npairs == len(want_pairs) and nman == int(hit.sum())


@pytest.mark.parametrize("per_wave", [1, 4])
def test_sphere_hull_per_wave_bitwise(pipe, oracle, per_wave):
    """Both GJK layouts of sphere-hull candidates (one per wavefront, or one
    per 16-lane row beside box-hull and hull-hull candidates) give the
    oracle's manifolds bit for bit, on a scene whose hull candidate count is
    not a multiple of four (so the last item has empty rows)."""
    sc = CASES["hulls4096"]
    print(f"sphere_hull_per_wave {per_wave}", flush=True)
    aabbs = oracle.compute_aabbs(sc)
    want_pairs, _ = oracle.find_pairs(aabbs, abi.UNIFORM_GRID)
    t = sc.shape_type
    hull = (t[want_pairs[:, 0]] == abi.HULL) | (t[want_pairs[:, 1]] == abi.HULL)
    sphere_hull = hull & ((t[want_pairs[:, 0]] == abi.SPHERE) | (t[want_pairs[:, 1]] == abi.SPHERE))
    assert hull.sum() % 4 != 0 and (hull & ~sphere_hull).sum() % 4 != 0
    pipe.set_colliders(sc)
    pipe.set_tuning(dict(sphere_hull_per_wave=per_wave))
    try:
        pipe.collide(algorithm=abi.UNIFORM_GRID)
        man = pipe.manifolds()
    finally:
        pipe.set_tuning(None)
    rec, hit = oracle.generate_contacts(sc, want_pairs)
    assert (hit & sphere_hull).sum() > 20
    assert_manifolds_equal(man, compact(rec, hit), f"sphere_hull_per_wave {per_wave}")


def test_boundary_bounds(pipe, oracle):
    """The reference's boundary AABB sets (tests/collision/broadphase_tests.cpp:89-105)."""
    F = np.float32
    sets = [
        np.zeros((0, 6), F),
        np.zeros((1, 6), F),
        np.tile(np.array([-1, -1, -1, 1, 1, 1], F), (16, 1)),
        np.array([[4, 0, 0, 5, 1, 1], [-2, -2, -2, 0, 0, 0], [0, 0, 0, 1, 1, 1],
                  [1, 0, 0, 2, 1, 1], [1, 1, 0, 2, 2, 1], [1, 1, 1, 2, 2, 2], [0, 4, 0, 1, 5, 1],
                  [0, 0, 4, 1, 1, 5], [1, 1, 1, 1, 1, 1], [-100, -0.5, -100, 100, 0, 100]], F),
    ]
    big = np.finfo(np.float32).max
    sets.append(np.array([[-0.5, -0.5, -0.5, 0.5, 0.5, 0.5], [0.5, -0.5, -0.5, 1.5, 0.5, 0.5],
                          [-1.5, -0.5, -0.5, -0.5, 0.5, 0.5],
                          [-100000, -0.5, -100000, 100000, 0, 100000],
                          [big, 0, 0, big, 0, 0], [big, 0, 0, big, 0, 0],
                          [-2147483648.0, 0, 0, -2147483648.0, 0, 0],
                          [-2147483648.0, 0, 0, -2147483648.0, 0, 0]], F))
    sets.append(np.array([[-1, -1, -1, 1, 1, 1], [-big, -big, -big, big, big, big]], F))
    for k, b in enumerate(sets):
        for algo in ALGOS:
            print(f"set {k} algo {algo}", flush=True)
            got = pipe.find_candidate_pairs(bounds=b, algorithm=algo)
            want, _ = oracle.find_pairs(b, algo)
            assert np.array_equal(got, want), (k, algo)


def _scene_from(g, prefix=""):
    get = (lambda k: g[prefix + k] if (prefix + k) in g.files else None)
    return scenes.Scene(pose=g[prefix + "pose"], shape_type=g[prefix + "shape_type"],
                        shape_data=g[prefix + "shape_data"], hull_range=get("hull_range"),
                        hull_points=get("hull_points"), body_position=get("body_position"),
                        local_pose=get("local_pose"), friction=get("friction"),
                        restitution=get("restitution"))


@pytest.mark.parametrize("name", ["spheres.npz", "mixed.npz"])
@pytest.mark.parametrize("algo", ALGOS)
def test_world_level_matches_reference(pipe, name, algo):
    """The collision section of the reference's PhysicsWorld::step (body-local
    anchors, combined friction/restitution, candidate order), recorded from the
    reference build, reproduced bit for bit."""
    from conftest import golden
    g = golden(name)
    sc = _scene_from(g, "w_")
    print(f"world {name} algo {algo}", flush=True)
    pipe.set_colliders(sc)
    pipe.set_world(sc)
    npairs, nman = pipe.collide(algorithm=algo)
    assert npairs == int(g["world_candidates"][0])
    man = pipe.manifolds()
    assert_manifolds_equal(man, g["world_contacts"], name + " world")
    st = pipe.narrowphase_stats()
    ws = g["world_np_stats"]
    assert [st["sphere_sphere_candidates"], st["sphere_box_candidates"], st["box_box_candidates"],
            st["sphere_sphere_contacts"], st["sphere_box_contacts"],
            st["box_box_contacts"]] == [int(v) for v in ws]


def test_world_same_body_pairs_dropped(pipe, oracle):
    """Colliders sharing a body never reach the narrow phase
    (physics_world.cpp:630-633): the manifolds are the one-body-per-collider
    result with those pairs removed, order kept."""
    sc = scenes.with_world_extras(scenes.mixed(3000, seed=31, side=scenes.mixed_side_for(3000, 2.0),
                                               with_hulls=False))
    want, _, _ = oracle.world_collide(sc)
    body = np.arange(sc.count, dtype=np.uint32) // 2  # colliders 2k and 2k+1 share body k
    sc.body = body
    pipe.set_colliders(sc)
    pipe.set_world(sc)
    pipe.collide(algorithm=abi.UNIFORM_GRID)
    man = pipe.manifolds()
    keep = body[want["first"]] != body[want["second"]]
    assert (~keep).sum() > 0
    assert_manifolds_equal(man, want[keep], "same-body filter")
