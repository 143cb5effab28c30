"""The reference grid's counters (findGridPairs, src/collision/broadphase.cpp:90-172:
cell size, entries, comparisons, overflow AABBs) reported by the HIP path's
broad-phase statistics, against the reference's own outputs held in the
golden fixtures and against the oracle on larger scenes and batches of
worlds.  Every call goes through the C ABI."""
import numpy as np
import pytest

from conftest import golden
from physgpu import abi, scenes

pytestmark = pytest.mark.gpu

GRID_FIELDS = ["grid_comparisons", "grid_entries", "grid_overflow_aabbs", "grid_cell_size"]
STAT_FIELDS = [k for k, _ in abi.BroadPhaseStats._fields_]


@pytest.fixture(scope="module")
def pipe(physgpu):
    p = physgpu.CollisionPipeline(70000, 1 << 21, max_hull_points=16 * 70000)
    yield p
    p.close()


def _grid(stats):
    return [stats[k] for k in GRID_FIELDS]


def _want(row):
    d = dict(zip(STAT_FIELDS, row))
    return [d[k] for k in GRID_FIELDS]


def test_bounds_fixtures(pipe):
    """All 128 bound sets of the fixture: the reference's tests' layouts
    (two boxes in 16 cells, the exact 64-cell limit, extreme coordinates,
    empty and single inputs) and seeded random sets."""
    g = golden("bounds.npz")
    names = sorted({k.split("__")[0] for k in g.files})
    print(f"{len(names)} bound sets", flush=True)
    for name in names:
        b = g[f"{name}__bounds"]
        pairs, st = pipe.find_candidate_pairs(b, abi.UNIFORM_GRID, with_stats=True)
        assert np.array_equal(pairs, g[f"{name}__pairs"]), name
        assert _grid(st) == _want(g[f"{name}__grid_stats"]), name
        assert st["aabb_pairs"] == len(pairs)


def test_reference_test_counters(pipe):
    """broadphase_tests.cpp:140-182: 16 entries / 8 comparisons / cell 1.0,
    the 64-cell boundary (1 overflow, 88 entries), 3 overflows at extremes."""
    g = golden("bounds.npz")
    _, st = pipe.find_candidate_pairs(g["grid_two__bounds"], abi.UNIFORM_GRID, with_stats=True)
    assert (st["grid_entries"], st["grid_comparisons"], st["grid_overflow_aabbs"],
            st["grid_cell_size"]) == (16, 8, 0, 1.0)
    _, st = pipe.find_candidate_pairs(g["grid_limit__bounds"], abi.UNIFORM_GRID, with_stats=True)
    assert (st["grid_overflow_aabbs"], st["grid_entries"]) == (1, 88)
    _, st = pipe.find_candidate_pairs(g["extremes__bounds"], abi.UNIFORM_GRID, with_stats=True)
    assert st["grid_overflow_aabbs"] == 3


@pytest.mark.parametrize("name", ["spheres.npz", "mixed.npz", "stack.npz"])
def test_scene_fixtures(pipe, name):
    g = golden(name)
    print(f"scene {name}", flush=True)
    _, st = pipe.find_candidate_pairs(g["aabbs"], abi.UNIFORM_GRID, with_stats=True)
    assert _grid(st) == _want(g["stats_grid"])


@pytest.mark.parametrize("n,seed", [(20000, 5), (65536, 6)])
def test_seeded_scenes_against_oracle(pipe, oracle, n, seed):
    sc = scenes.mixed(n, seed=seed, side=scenes.mixed_side_for(n, 1.0), with_hulls=False)
    print(f"mixed {n} seed {seed}", flush=True)
    pipe.set_colliders(sc)
    aabbs = pipe.compute_aabbs()
    _, st = pipe.find_candidate_pairs(algorithm=abi.UNIFORM_GRID, with_stats=True)
    _, want = oracle.find_pairs(aabbs, abi.UNIFORM_GRID)
    assert _grid(st) == _grid(want)
    assert st["grid_comparisons"] > st["aabb_pairs"] > 0


def test_forced_hash_collisions(pipe, oracle, monkeypatch):
    """Cell tags kept to 16 bits: hundreds of cells collide in the table and
    are counted in further rounds; the counters do not change."""
    sc = scenes.mixed(12000, seed=8, side=scenes.mixed_side_for(12000, 1.0), with_hulls=False)
    pipe.set_colliders(sc)
    aabbs = pipe.compute_aabbs()
    _, want = oracle.find_pairs(aabbs, abi.UNIFORM_GRID)
    assert want["grid_entries"] > 8000
    monkeypatch.setenv("PHYSGPU_GRID_STATS_TAG_BITS", "16")
    print("tag bits 16", flush=True)
    _, st = pipe.find_candidate_pairs(algorithm=abi.UNIFORM_GRID, with_stats=True)
    assert _grid(st) == _grid(want)


def test_batch_of_worlds(pipe, oracle):
    """Each world is its own reference grid: counts add up over the worlds,
    the cell reported is the largest world cell."""
    parts = [scenes.mixed(n, seed=40 + n, side=scenes.mixed_side_for(n, s), with_hulls=False)
             for n, s in ((17, 1.0), (300, 0.5), (1000, 2.0), (64, 1.0), (1, 1.0))]
    sc, off = scenes.concat(parts)
    pipe.set_colliders(sc)
    aabbs = pipe.compute_aabbs()
    pipe.set_worlds(off)
    print(f"batch of {len(parts)} worlds", flush=True)
    pipe.collide(algorithm=abi.UNIFORM_GRID)
    st = pipe.step_stats()["broad_phase"]
    pipe.set_worlds(None)
    want = [oracle.find_pairs(aabbs[off[w]:off[w + 1]], abi.UNIFORM_GRID)[1]
            for w in range(len(parts))]
    for k in ("grid_comparisons", "grid_entries", "grid_overflow_aabbs"):
        assert st[k] == sum(s[k] for s in want), k
    assert st["grid_cell_size"] == max(s["grid_cell_size"] for s in want)
