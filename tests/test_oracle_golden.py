"""The CPU oracle reproduces the reference engine's own outputs, bit for bit,
on the committed golden fixtures (written by tests/golden/make_golden.py from
the reference compiled from its sources)."""
import numpy as np
import pytest

from conftest import assert_manifolds_equal, golden
from physgpu import abi, scenes

STAT_FIELDS = [k for k, _ in abi.BroadPhaseStats._fields_]
TIMING = {"record_build_ms", "record_sort_ms", "sweep_ms", "pair_sort_ms"}


def scene_from(g, prefix=""):
    get = (lambda k: g[prefix + k] if (prefix + k) in g.files else None)
    return scenes.Scene(pose=g[prefix + "pose"], shape_type=g[prefix + "shape_type"],
                        shape_data=g[prefix + "shape_data"], hull_range=get("hull_range"),
                        hull_points=get("hull_points"), body_position=get("body_position"),
                        local_pose=get("local_pose"), friction=get("friction"),
                        restitution=get("restitution"), body=get("body"))


def stats_match(got, want_row, algo):
    for k, v in zip(STAT_FIELDS, want_row):
        if k in TIMING:
            continue
        if algo == 2 and k.startswith("tree_") and k != "tree_proxy_count":
            continue  # tree work counters describe the reference's dynamic tree
        assert got[k] == v, (k, got[k], v)


@pytest.mark.parametrize("name", ["spheres.npz", "mixed.npz", "stack.npz"])
def test_scene_fixture(oracle, name):
    g = golden(name)
    sc = scene_from(g)
    aabbs = oracle.compute_aabbs(sc)
    assert np.array_equal(aabbs.view(np.uint32), g["aabbs"].view(np.uint32))
    for algo, an in ((0, "sap"), (1, "grid"), (2, "tree")):
        pairs, st = oracle.find_pairs(g["aabbs"], algo)
        assert np.array_equal(pairs, g[f"pairs_{an}"])
        stats_match(st, g[f"stats_{an}"], algo)
    rec, hit = oracle.generate_contacts(sc, g["pairs_sap"])
    assert np.array_equal(hit, g["hit"])
    assert_manifolds_equal(rec[hit], g["contacts"][g["hit"]], name)


@pytest.mark.parametrize("name", ["spheres.npz", "mixed.npz"])
def test_world_fixture(oracle, name):
    g = golden(name)
    sc = scene_from(g, "w_")
    m, t, s = oracle.world_collide(sc)
    assert t["candidate_pairs"] == int(g["world_candidates"][0])
    assert np.array_equal(s, g["world_np_stats"])
    assert_manifolds_equal(m, g["world_contacts"], name + " world")


def test_kat_fixture(oracle):
    g = golden("kat.npz")
    rec, hit = oracle.generate_contacts(scene_from(g), g["pairs"])
    assert np.array_equal(hit, g["hit"]) and hit.all()
    assert_manifolds_equal(rec, g["contacts"], "kat")


def test_bounds_fixture(oracle):
    g = golden("bounds.npz")
    names = sorted({k.split("__")[0] for k in g.files})
    assert len(names) == 8 + 120
    for name in names:
        b = g[f"{name}__bounds"]
        for algo, an in ((0, "sap"), (1, "grid"), (2, "tree")):
            pairs, st = oracle.find_pairs(b, algo)
            assert np.array_equal(pairs, g[f"{name}__pairs"]), (name, an)
            stats_match(st, g[f"{name}__{an}_stats"], algo)


def test_grid_counters_of_reference_tests(oracle):
    """broadphase_tests.cpp:140-182: 16 entries / 8 comparisons / cell 1.0,
    the exact 64-cell boundary (1 overflow, 88 entries), 3 overflows at extremes."""
    g = golden("bounds.npz")
    _, st = oracle.find_pairs(g["grid_two__bounds"], abi.UNIFORM_GRID)
    assert (st["grid_entries"], st["grid_comparisons"], st["grid_overflow_aabbs"],
            st["grid_cell_size"]) == (16, 8, 0, 1.0)
    _, st = oracle.find_pairs(g["grid_limit__bounds"], abi.UNIFORM_GRID)
    assert (st["grid_overflow_aabbs"], st["grid_entries"]) == (1, 88)
    _, st = oracle.find_pairs(g["extremes__bounds"], abi.UNIFORM_GRID)
    assert st["grid_overflow_aabbs"] == 3


@pytest.mark.parametrize("defaults", [False, True])
def test_gjk_epa_fixture(oracle, defaults):
    g = golden("gjk.npz")
    sc = scene_from(g)
    p = "d_" if defaults else ""
    it, tol = (32, 1e-4) if defaults else (64, 1e-6)
    st, nrm, dep, smp = oracle.gjk_epa(sc, g["pairs"], it, 64, tol)
    assert np.array_equal(st, g[p + "status"])
    assert np.array_equal(nrm.view(np.uint32), g[p + "normal"].view(np.uint32))
    assert np.array_equal(dep.view(np.uint32), g[p + "depth"].view(np.uint32))
    assert np.array_equal(smp.view(np.uint32), g[p + "simplex"].view(np.uint32))
    assert (st == 1).sum() > 100 and (st == 0).sum() > 100


def test_bounded_epa_equals_reference_within_bounds(oracle):
    """With the GPU's polytope bounds and no continuation the oracle's EPA is
    the reference's on every pair that stays inside them; the others report
    overflow (status 3)."""
    g = golden("gjk.npz")
    sc = scene_from(g)
    oracle.mirror_gpu_epa(continuation=False)
    try:
        st, nrm, dep, _ = oracle.gjk_epa(sc, g["pairs"], 64, 64, 1e-6)
    finally:
        oracle.mirror_gpu_epa(False)
    inside = st != 3
    assert inside.mean() > 0.97 and (~inside).sum() > 0
    assert np.array_equal(st[inside], g["status"][inside])
    assert np.array_equal(nrm[inside].view(np.uint32), g["normal"][inside].view(np.uint32))
    assert np.array_equal(dep[inside].view(np.uint32), g["depth"][inside].view(np.uint32))
    assert np.all(g["status"][~inside] >= 1)  # GJK hits the reference kept expanding


def test_gpu_epa_equals_reference(oracle):
    """The GPU's EPA (the reference's list within the LDS bounds, continued
    over distinct faces with copy counts past them) gives the reference's
    status, normal and depth bits on every pair of the fixture, the pairs
    that outgrow the bounds included."""
    g = golden("gjk.npz")
    sc = scene_from(g)
    oracle.mirror_gpu_epa()
    try:
        st, nrm, dep, _ = oracle.gjk_epa(sc, g["pairs"], 64, 64, 1e-6)
    finally:
        oracle.mirror_gpu_epa(False)
    assert not np.any(st == 3)
    assert np.array_equal(st, g["status"])
    assert np.array_equal(nrm.view(np.uint32), g["normal"].view(np.uint32))
    assert np.array_equal(dep.view(np.uint32), g["depth"].view(np.uint32))


@pytest.mark.parametrize("algo", [0, 1, 2])
def test_world_compound_bodies_match_reference(oracle, algo):
    """PhysicsWorld::step with 1 to 4 colliders per body (world_bodies.npz,
    recorded from the reference): same-body pairs erased in order
    (physics_world.cpp:630-633), body-local anchors, combined materials,
    StepStats candidatePairs / possiblePairs / contactCount and the narrow
    phase counters, bit for bit (the oracle's first / second are collider
    indices, the reference's body indices)."""
    g = golden("world_bodies.npz")
    sc = scene_from(g)
    name = ("sap", "grid", "tree")[algo]
    m, t, st = oracle.world_collide(sc, algorithm=algo)
    want = g["contacts_" + name]
    got = m.copy()
    got["first"] = sc.body[m["first"]]
    got["second"] = sc.body[m["second"]]
    w = lambda a: np.ascontiguousarray(a).view(np.uint32).reshape(len(a), 40)[:, :36]  # noqa
    assert w(got).shape == w(want).shape and np.array_equal(w(got), w(want))
    cand, possible, contacts = (int(v) for v in g["counts_" + name])
    assert (t["candidate_pairs"], t["possible_pairs"], len(m)) == (cand, possible, contacts)
    assert np.array_equal(st, g["np_stats_" + name])
    # possiblePairs in closed form: C(n, 2) minus the pairs inside each body.
    k = np.bincount(sc.body)
    assert possible == sc.count * (sc.count - 1) // 2 - int((k * (k - 1) // 2).sum())
    # The filter removed something, so the test exercises it.
    assert cand < oracle.find_pairs(oracle.compute_aabbs(sc), algo)[1]["aabb_pairs"]
