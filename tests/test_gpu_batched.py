"""Batches of worlds in one context (one launch per phase), the HIP-graph form
of the step and the autotuner.  Every world of a batch must produce exactly the
pairs and manifolds it produces alone (the oracle on that world; the per-world
semantics of PhysicsWorld::step, physics_world.cpp:518-782), and tuning or
graph replay must not change a bit."""
import warnings

import numpy as np
import pytest

from conftest import GEOM_WORDS, words
from physgpu import abi, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def bounded_oracle(oracle):
    oracle.mirror_gpu_epa()
    yield
    oracle.mirror_gpu_epa(False)


@pytest.fixture(scope="module")
def batch():
    return scenes.batched_worlds(96, seed=41, min_bodies=16, max_bodies=1024)


@pytest.fixture(scope="module")
def bpipe(physgpu, batch):
    sc, off, _ = batch
    p = physgpu.CollisionPipeline(sc.count, 8 * sc.count,
                                  max_hull_points=sc.hull_points.shape[1])
    yield p
    p.close()


def _world_expect(oracle, sc, off, w):
    sub = sc.subset(np.arange(off[w], off[w + 1]))
    pairs, _ = oracle.find_pairs(oracle.compute_aabbs(sub), abi.UNIFORM_GRID)
    rec, hit = oracle.generate_contacts(sub, pairs)
    return pairs, rec[np.asarray(hit, bool)], np.nonzero(hit)[0]


def _check_world(pairs_g, man_g, po, mo, off, w, want_pairs, want_man, want_idx):
    p = pairs_g[po[w]:po[w + 1]].astype(np.int64) - int(off[w])
    assert np.array_equal(p, want_pairs.astype(np.int64)), f"world {w}: pair sets differ"
    m = man_g[mo[w]:mo[w + 1]]
    assert len(m) == len(want_man), f"world {w}: {len(m)} manifolds vs {len(want_man)}"
    if not len(m):
        return
    g, e = words(m), words(want_man)
    assert np.array_equal(g[:, 0:2].astype(np.int64) - int(off[w]), e[:, 0:2].astype(np.int64))
    bad = np.nonzero((g[:, GEOM_WORDS] != e[:, GEOM_WORDS]).any(axis=1))[0]
    assert not len(bad), f"world {w}: {len(bad)} manifolds differ bitwise"
    assert np.array_equal(m["pair_index"].astype(np.int64) - int(po[w]), want_idx)


def test_batch_equals_each_world_alone(bpipe, batch, oracle):
    sc, off, _ = batch
    print(f"batch of {len(off) - 1} worlds, {sc.count} bodies", flush=True)
    bpipe.set_colliders(sc)
    bpipe.set_worlds(off)
    npairs, nman = bpipe.collide(algorithm=abi.UNIFORM_GRID)
    pairs = bpipe.pairs()
    man = bpipe.manifolds()
    po, mo = bpipe.world_ranges()
    assert po[-1] == npairs == len(pairs) and mo[-1] == nman == len(man)
    assert npairs > sc.count // 2
    for w in range(len(off) - 1):
        _check_world(pairs, man, po, mo, off, w, *_world_expect(oracle, sc, off, w))


def test_full_batch_largest_and_every_64th_world(physgpu, oracle):
    """The full batch of 4,096 uneven worlds (16 to 4,096 bodies, about 3M
    bodies, where per-world key hashes share the cell table most densely):
    the largest world and every 64th world against the oracle run on that
    world alone, and the batch's totals against the per-world ranges."""
    sc, off, counts = scenes.batched_worlds(4096)
    print(f"full batch: {len(off) - 1} worlds, {sc.count} bodies", flush=True)
    with physgpu.CollisionPipeline(sc.count, 4 * sc.count,
                                   max_hull_points=sc.hull_points.shape[1]) as p:
        p.set_colliders(sc)
        p.set_worlds(off)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", physgpu.EpaUnresolvedWarning)
            npairs, nman = p.collide_graph(algorithm=abi.UNIFORM_GRID)
            pairs, man = p.pairs(), p.manifolds()
            po, mo = p.world_ranges()
            unresolved = p.unresolved_pairs()
    assert po[-1] == npairs == len(pairs) and mo[-1] == nman == len(man)
    assert npairs > sc.count // 2
    # Pairs the GPU leaves without an EPA result: the oracle with the same
    # polytope bounds cannot decide them either (3: past the distinct-face
    # bound, 4: tied closest faces of unknown order).
    if len(unresolved):
        st, _, _, _ = oracle.gjk_epa(sc, unresolved)
        print(f"unresolved hull pairs {unresolved.tolist()} oracle status {st.tolist()}",
              flush=True)
        assert np.isin(st, (3, 4)).all()
    worlds = sorted(set(range(0, len(off) - 1, 64)) | {int(np.argmax(counts))})
    for w in worlds:
        _check_world(pairs, man, po, mo, off, w, *_world_expect(oracle, sc, off, w))


def test_batch_rejects_other_broadphases(bpipe, batch):
    sc, off, _ = batch
    bpipe.set_colliders(sc)
    bpipe.set_worlds(off)
    from physgpu import PhysGpuError
    for algo in (abi.SWEEP_AND_PRUNE, abi.DYNAMIC_TREE):
        with pytest.raises(PhysGpuError):
            bpipe.collide(algorithm=algo)


def test_set_worlds_validates_offsets(bpipe, batch):
    sc, off, _ = batch
    bpipe.set_colliders(sc)
    bad = off.copy()
    bad[-1] -= 1
    with pytest.raises(ValueError):
        bpipe.set_worlds(bad)
    bad = off.copy()
    bad[3], bad[4] = bad[4], bad[3]
    with pytest.raises(ValueError):
        bpipe.set_worlds(bad)


def test_graph_replays_match_eager(bpipe, batch):
    """Replays of the captured step, with eager steps and host reads between
    them, and with poses changed in place between replays."""
    sc, off, _ = batch
    bpipe.set_colliders(sc)
    bpipe.set_worlds(off)
    bpipe.collide(algorithm=abi.UNIFORM_GRID)
    want_p, want_m = bpipe.pairs(), bpipe.manifolds()
    want_r = bpipe.world_ranges()
    rng = np.random.default_rng(3)
    moved = sc.pose.copy()
    moved[:3] += rng.uniform(-0.05, 0.05, moved[:3].shape).astype(np.float32)
    bpipe.update_poses(moved)
    bpipe.collide(algorithm=abi.UNIFORM_GRID)
    moved_p, moved_m = bpipe.pairs(), bpipe.manifolds()
    bpipe.update_poses(sc.pose)
    for replay in range(4):
        print(f"replay {replay}", flush=True)
        bpipe.collide_graph(algorithm=abi.UNIFORM_GRID)
        assert np.array_equal(bpipe.pairs(), want_p)
        assert np.array_equal(words(bpipe.manifolds()), words(want_m))
        r = bpipe.world_ranges()
        assert np.array_equal(r[0], want_r[0]) and np.array_equal(r[1], want_r[1])
        bpipe.collide(algorithm=abi.UNIFORM_GRID)  # eager work between replays
    bpipe.update_poses(moved)
    bpipe.collide_graph(algorithm=abi.UNIFORM_GRID)
    assert np.array_equal(bpipe.pairs(), moved_p)
    assert np.array_equal(words(bpipe.manifolds()), words(moved_m))
    bpipe.update_poses(sc.pose)


def test_graph_single_world(physgpu, oracle):
    sc = scenes.mixed(4096, seed=43, side=scenes.mixed_side_for(4096, 1.0))
    with physgpu.CollisionPipeline(4096, 1 << 16, max_hull_points=sc.hull_points.shape[1]) as p:
        p.set_colliders(sc)
        for algo in (abi.SWEEP_AND_PRUNE, abi.UNIFORM_GRID, abi.DYNAMIC_TREE):
            print(f"graph algo {algo}", flush=True)
            p.collide(algorithm=algo)
            want_p, want_m = p.pairs(), p.manifolds()
            for _ in range(3):
                p.collide_graph(algorithm=algo)
                assert np.array_equal(p.pairs(), want_p)
                assert np.array_equal(words(p.manifolds()), words(want_m))
        pairs, _ = oracle.find_pairs(oracle.compute_aabbs(sc))
        assert np.array_equal(want_p, pairs)


def test_autotune_keeps_results(bpipe, batch):
    sc, off, _ = batch
    bpipe.set_colliders(sc)
    bpipe.set_worlds(off)
    bpipe.set_tuning(None)
    bpipe.collide(algorithm=abi.UNIFORM_GRID)
    want_p, want_m = bpipe.pairs(), bpipe.manifolds()
    t = bpipe.autotune(algorithm=abi.UNIFORM_GRID, repeats=1)
    print(f"tuning {t}", flush=True)
    assert t["sort_items"] in (4, 8, 16)
    assert all(b in (64, 128, 256) for b in t["class_block"])
    assert bpipe.tuning() == t
    bpipe.collide(algorithm=abi.UNIFORM_GRID)
    assert np.array_equal(bpipe.pairs(), want_p)
    assert np.array_equal(words(bpipe.manifolds()), words(want_m))
    assert t["sphere_hull_per_wave"] in (1, 4)
    for items, epa, shw in ((4, 7, 4), (16, 10000, 1)):
        bpipe.set_tuning(dict(t, sort_items=items, class_block=[64, 128, 64],
                              class_groups=[7, 300, 5000], query_block=64, query_groups=33,
                              epa_groups=epa, sphere_hull_per_wave=shw))
        bpipe.collide_graph(algorithm=abi.UNIFORM_GRID)
        assert np.array_equal(bpipe.pairs(), want_p)
        assert np.array_equal(words(bpipe.manifolds()), words(want_m))
    with pytest.raises(ValueError):
        bpipe.set_tuning(dict(sort_items=5))
    with pytest.raises(ValueError):
        bpipe.set_tuning(dict(sphere_hull_per_wave=2))
    bpipe.set_tuning(None)


def test_per_world_loop_matches_batch(physgpu, batch):
    """The plain pipeline run world by world (what the batch replaces) gives the
    batch's per-world results."""
    sc, off, _ = batch
    with physgpu.CollisionPipeline(sc.count, 8 * sc.count,
                                   max_hull_points=sc.hull_points.shape[1]) as b:
        b.set_colliders(sc)
        b.set_worlds(off)
        b.collide_graph(algorithm=abi.UNIFORM_GRID)
        pairs, man = b.pairs(), b.manifolds()
        po, mo = b.world_ranges()
    with physgpu.CollisionPipeline(4096, 1 << 16, max_hull_points=4096 * 16) as one:
        for w in range(0, len(off) - 1, 7):
            sub = sc.subset(np.arange(off[w], off[w + 1]))
            one.set_colliders(sub)
            one.collide(algorithm=abi.UNIFORM_GRID)
            wm = one.manifolds()
            _check_world(pairs, man, po, mo, off, w, one.pairs(), wm,
                         wm["pair_index"].astype(np.int64))
