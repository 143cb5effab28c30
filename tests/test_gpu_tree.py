"""Persistent proxies on the device (physgpu_tree_*), in the shape of the
reference's tests/collision/dynamic_aabb_tree_tests.cpp: fat bounds and
resizing (:63-91), a balanced lifecycle of 512 proxies over 80 frames of
moves, destructions and re-creations checked against brute force (:93-156),
proxy reuse and full-width user indices (:158-195), invalid calls (:197-226).
Expected pairs are the closed form the reference's matchesBruteForce uses
(:24-60): every pair of live proxies whose tight bounds overlap inclusively,
as (smaller, larger) user index, sorted.  The refit policy must return the
rebuild policy's pairs."""
import numpy as np
import pytest

from physgpu import abi

pytestmark = pytest.mark.gpu


def brute_force(proxies):
    live = [(b, u) for (pid, b, u) in proxies if pid is not None]
    out = []
    for i in range(len(live)):
        for j in range(i + 1, len(live)):
            a, b = live[i][0], live[j][0]
            if np.all(a[:3] <= b[3:]) and np.all(a[3:] >= b[:3]):
                u, v = live[i][1], live[j][1]
                out.append((min(u, v), max(u, v)))
    out.sort()
    return np.array(out, np.uint64).reshape(-1, 2)


def matches(tree, proxies):
    want = brute_force(proxies)
    stats = {}
    got = tree.find_candidate_pairs(stats)
    live = sum(pid is not None for pid, _, _ in proxies)
    assert np.array_equal(got, want), f"pairs {got.tolist()} != {want.tolist()}"
    assert tree.size() == live and stats["tree_proxy_count"] == live
    assert stats["aabb_pairs"] == len(want) and stats["tree_leaf_checks"] >= len(want)
    if live == 0:
        assert stats["tree_node_pair_visits"] == 0 and stats["tree_height"] == 0
    return True


def box(lo, hi):
    return np.array(list(lo) + list(hi), np.float32)


@pytest.fixture(params=["rebuild", "refit"])
def tree(physgpu, request):
    t = physgpu.DynamicAabbTree(capacity=4096, max_pairs=1 << 16, policy=request.param)
    yield t
    t.close()


def test_fat_bounds_and_resizing(tree):
    p = [[None, box((0, 0, 0), (1, 1, 1)), 7], [None, box((1.05, 0, 0), (2.05, 1, 1)), 2]]
    for q in p:
        q[0] = tree.create_proxy(q[1], q[2])
    assert matches(tree, p) and len(tree.find_candidate_pairs()) == 0
    p[1][1] = box((0.98, 0, 0), (1.98, 1, 1))
    assert not tree.update_proxy(p[1][0], p[1][1])  # inside the fat bounds
    assert matches(tree, p) and len(tree.find_candidate_pairs()) == 1
    p[1][1] = box((10, 0, 0), (11, 1, 1))
    assert tree.update_proxy(p[1][0], p[1][1]) and matches(tree, p)
    p[1][1] = box((0, 0, 0), (100, 100, 100))
    assert tree.update_proxy(p[1][0], p[1][1]) and matches(tree, p)
    p[1][1] = box((10, 0, 0), (11, 1, 1))
    # Oversized fat bounds shrink (reinsert), then a stationary update keeps them.
    assert tree.update_proxy(p[1][0], p[1][1])
    assert not tree.update_proxy(p[1][0], p[1][1])
    assert matches(tree, p)


def test_balanced_lifecycle(tree):
    p = []
    for i in range(512):
        x = np.float32(i) * np.float32(0.7)
        b = box((x, 0, 0), (x + np.float32(1), 1, 1))
        u = (511 - i) * 17
        p.append([tree.create_proxy(b, u), b, u])
    assert matches(tree, p)
    rng = np.random.default_rng(123)
    for frame in range(80):
        moved_ids, moved_bounds = [], []
        for i, q in enumerate(p):
            if (i + frame) % 19 == 0:
                if q[0] is not None:
                    tree.destroy_proxy(q[0])
                    q[0] = None
                else:
                    q[0] = tree.create_proxy(q[1], q[2])
            if q[0] is None:
                continue
            c = rng.uniform(-12.0, 12.0, 3).astype(np.float32)
            e = rng.uniform(0.01, 2.0, 3).astype(np.float32)
            q[1] = np.concatenate([c - e, c + e]).astype(np.float32)
            moved_ids.append(q[0])
            moved_bounds.append(q[1])
        tree.update_proxies(moved_ids, moved_bounds)
        assert matches(tree, p)
        if frame % 10 == 0:
            # Moves only (no proxy created or destroyed): the refit policy
            # keeps the build and must agree with brute force.
            ids = [q[0] for q in p if q[0] is not None]
            for q in p:
                if q[0] is not None:
                    q[1] = (q[1] + np.float32(0.05)).astype(np.float32)
            tree.update_proxies(ids, [q[1] for q in p if q[0] is not None])
            assert matches(tree, p)
    for q in p:
        if q[0] is not None:
            tree.destroy_proxy(q[0])
            q[0] = None
    assert matches(tree, p)
    p[0][0] = tree.create_proxy(p[0][1], p[0][2])
    assert matches(tree, p)


def test_proxy_reuse_and_wide_ids(tree):
    wide = (1 << 64) - 2
    p = [[None, box((0, 0, 0), (1, 1, 1)), 10], [None, box((10, 0, 0), (11, 1, 1)), 20],
         [None, box((0.5, 0, 0), (1.5, 1, 1)), wide]]
    for q in p:
        q[0] = tree.create_proxy(q[1], q[2])
    assert matches(tree, p)
    tree.destroy_proxy(p[0][0])
    p[0][0] = None
    p[2][1] = box((0.55, 0, 0), (1.55, 1, 1))
    assert not tree.update_proxy(p[2][0], p[2][1])
    assert matches(tree, p)
    p[0][1] = box((10.5, 0, 0), (11.5, 1, 1))
    p[0][0] = tree.create_proxy(p[0][1], p[0][2])
    tree.destroy_proxy(p[1][0])
    p[1][1] = box((0.6, 0, 0), (1.6, 1, 1))
    p[1][0] = tree.create_proxy(p[1][1], p[1][2])
    assert matches(tree, p)
    got = tree.find_candidate_pairs()
    assert int(got[-1][1]) == wide
    # A proxy id wider than 32 bits is invalid, never an alias of a live one.
    with pytest.raises(ValueError):
        tree.update_proxy((1 << 32) + p[0][0], box((50, 50, 50), (51, 51, 51)))
    assert matches(tree, p)


def test_invalid_calls(tree):
    with pytest.raises(ValueError):
        tree.create_proxy(box((1, 0, 0), (0, 1, 1)), 0)
    pid = tree.create_proxy(box((0, 0, 0), (1, 1, 1)), 0)
    with pytest.raises(ValueError):
        tree.update_proxy(pid, box((np.nan, 0, 0), (1, 1, 1)))
    with pytest.raises(ValueError):
        tree.create_proxy(box((0, 0, 0), (1, 1, 1)), 0)  # user index held by a live proxy
    tree.destroy_proxy(pid)
    with pytest.raises(ValueError):
        tree.destroy_proxy(pid)
    assert tree.size() == 0 and len(tree.find_candidate_pairs()) == 0


def test_many_proxies_match_oracle(physgpu, oracle):
    """20,000 random proxies with scattered user indices, moved twice: both
    policies give the oracle's pair set mapped to user indices."""
    rng = np.random.default_rng(7)
    n = 20000
    c = rng.uniform(-40, 40, (n, 3)).astype(np.float32)
    e = rng.uniform(0.1, 1.2, (n, 3)).astype(np.float32)
    users = rng.permutation(10 * n)[:n].astype(np.uint64) * np.uint64(3) + np.uint64(1)
    for policy in ("rebuild", "refit"):
        t = physgpu.DynamicAabbTree(capacity=n, max_pairs=1 << 20, policy=policy)
        try:
            bounds = np.concatenate([c - e, c + e], axis=1).astype(np.float32)
            ids = t.create_proxies(bounds, users)
            for step in range(3):
                want_idx, _ = oracle.find_pairs(bounds.reshape(-1, 6), abi.SWEEP_AND_PRUNE)
                mapped = np.sort(users[want_idx.astype(np.int64)], axis=1)
                want = mapped[np.lexsort((mapped[:, 1], mapped[:, 0]))]
                got = t.find_candidate_pairs()
                print(f"policy {policy} step {step}: {len(got)} pairs", flush=True)
                assert np.array_equal(got, want)
                bounds = bounds + rng.uniform(-0.3, 0.3, (n, 1)).astype(np.float32)
                t.update_proxies(ids, bounds)
        finally:
            t.close()
