"""Full-size workloads on the GPU: 1,048,576 mixed bodies (spheres, boxes,
16-point hulls) and the 262,145-box resting stack.

Pair sets are compared with the SHA-256 of the reference engine's own pair
list (tests/golden/large.json, written from the reference build); manifolds
bit for bit with the CPU oracle on the same pairs, plus the counts the
reference produced for the stack."""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, assert_manifolds_equal, compact
from physgpu import abi, scenes

pytestmark = pytest.mark.gpu

LARGE = json.load(open(os.path.join(GOLDEN, "large.json")))
ALGOS = [abi.SWEEP_AND_PRUNE, abi.UNIFORM_GRID, abi.DYNAMIC_TREE]


def sha(pairs):
    return hashlib.sha256(np.ascontiguousarray(pairs, np.uint32).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def mixed_1m():
    info = LARGE["mixed1M"]
    return scenes.mixed(info["n"], seed=info["seed"], side=info["side"])


@pytest.fixture(scope="module")
def big_pipe(physgpu):
    p = physgpu.CollisionPipeline(1 << 20, 1 << 20, max_hull_points=1 << 21)
    yield p
    p.close()


@pytest.fixture(scope="module", autouse=True)
def bounded_oracle(oracle):
    oracle.mirror_gpu_epa()
    yield
    oracle.mirror_gpu_epa(False)


@pytest.mark.parametrize("algo", ALGOS)
def test_mixed_1m_pairs_match_reference(big_pipe, mixed_1m, algo):
    print(f"mixed 1M algo {algo}", flush=True)
    big_pipe.set_colliders(mixed_1m)
    npairs, _ = big_pipe.collide(algorithm=algo)
    pairs = big_pipe.pairs()
    info = LARGE["mixed1M"]
    assert npairs == info["pairs"] == len(pairs)
    assert sha(pairs) == info["sha256"]


def test_mixed_1m_manifolds_bitwise(big_pipe, mixed_1m, oracle):
    print("mixed 1M manifolds", flush=True)
    big_pipe.set_colliders(mixed_1m)
    big_pipe.collide(algorithm=abi.UNIFORM_GRID)
    pairs = big_pipe.pairs()
    man = big_pipe.manifolds()
    stats = big_pipe.narrowphase_stats()
    rec, hit = oracle.generate_contacts(mixed_1m, pairs, workers=min(os.cpu_count() or 1, 16))
    assert_manifolds_equal(man, compact(rec, hit), "mixed 1M")
    hull = ((mixed_1m.shape_type[pairs[:, 0]] == abi.HULL) |
            (mixed_1m.shape_type[pairs[:, 1]] == abi.HULL))
    st, _, _, _ = oracle.gjk_epa(mixed_1m, pairs[hull])
    assert stats["epa_unresolved"] == int((st == 3).sum()) == 0
    assert stats["gjk_epa_failures"] == int((st == 2).sum())
    # Every pair whose polytope outgrows the LDS bounds is in the fixture of
    # the reference's own unbounded results, and was continued on the GPU.
    ovf = np.load(os.path.join(GOLDEN, "epa_overflow.npz"), allow_pickle=False)
    assert stats["epa_overflows"] == len(ovf["pairs"])
    check_against_reference(man, ovf)


def check_against_reference(man, ovf):
    """GPU manifolds of the fixture's pairs against the reference's unbounded
    computeEpa (tests/golden/make_epa_overflow.py): a contact with the
    reference's normal and depth bits where it converged, none where it
    failed (pairs the reference did not finish within its limits are only
    counted)."""
    got = {(int(m["first"]), int(m["second"])): m for m in man
           if m["shape_pair"] >= abi.PAIR_SPHERE_HULL}
    converged = 0
    for (a, b), rs, rn, rd in zip(ovf["pairs"], ovf["ref_status"], ovf["ref_normal"],
                                  ovf["ref_depth"]):
        m = got.get((int(a), int(b)))
        if rs == 1:
            assert m is not None, (a, b)
            assert np.array_equal(m["normal"].view(np.uint32), rn.view(np.uint32)), (a, b)
            assert m["penetration"][0].view(np.uint32) == rd.view(np.uint32), (a, b)
            converged += 1
        elif rs == 2:
            assert m is None, (a, b)
    assert converged == int((ovf["ref_status"] == 1).sum()) > 0


def test_stack_262145(big_pipe, oracle):
    info = LARGE["stack262145"]
    stk = scenes.stack(columns=64, height=64, seed=info["seed"])
    big_pipe.set_colliders(stk)
    for algo in ALGOS:
        print(f"stack algo {algo}", flush=True)
        npairs, nman = big_pipe.collide(algorithm=algo)
        pairs = big_pipe.pairs()
        assert npairs == info["pairs"] and sha(pairs) == info["sha256"]
        assert nman == info["manifolds"]
    man = big_pipe.manifolds()
    assert int(man["point_count"].sum()) == info["points"]
    rec, hit = oracle.generate_contacts(stk, pairs, workers=min(os.cpu_count() or 1, 16))
    assert_manifolds_equal(man, compact(rec, hit), "stack")
