"""Write epa_overflow.npz: the reference's own unbounded EPA result on every
hull pair of the 1,048,576-body mixed scene whose polytope outgrows the GPU's
bounds (PHYSGPU_EPA_MAX_FACES faces / PHYSGPU_EPA_MAX_HORIZON horizon edges).

Runs only where the reference sources exist (/root/reference): the reference
shim oracle/_ref/libphysref.so (the reference's intersectGjk + computeEpa,
src/collision/gjk.cpp:128-158 and epa.cpp:64-130, compiled from its sources
with this project's support maps) is called on each such pair in a child
process of its own, with a wall-clock and memory limit, because the FP32
polytope can grow without bound there (the horizon dedup of epa.cpp:49-60 is
quadratic in its length).

The fixture holds data only: the two colliders of each pair (pose, shape,
hull points), their indices in the 1M scene, the bounded oracle's status and
the reference's outcome:
  ref_status  0 GJK miss, 1 EPA converged, 2 EPA failure (horizon empty or
              iteration cap), -1 not finished within the limits below
  ref_normal, ref_depth   the converged result (zero otherwise)
  ref_seconds            wall time of the reference call
"""
import concurrent.futures as cf
import json
import multiprocessing as mp
import os
import resource
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "phys-engine_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from physgpu import abi, scenes  # noqa: E402

SEED, N, SIDE = 20261203, 1048576, 256.0
TIME_LIMIT_S = float(os.environ.get("EPA_REF_TIME_LIMIT", "1800"))
MEMORY_LIMIT_B = 6 << 30
# Optional JSON-lines file of finished pairs (EPA_REF_CACHE): a run that is
# interrupted and started again skips the pairs already in it.
CACHE = os.environ.get("EPA_REF_CACHE")


def overflow_pairs():
    """Hull pairs of the 1M scene whose bounded oracle EPA overflows."""
    sc = scenes.mixed(N, seed=SEED, side=SIDE)
    orc = O.CpuCollision("oracle")
    aabbs = orc.compute_aabbs(sc)
    pairs, _ = orc.find_pairs(aabbs, abi.UNIFORM_GRID)
    hull = sc.shape_type == abi.HULL
    hp = pairs[hull[pairs[:, 0]] | hull[pairs[:, 1]]]
    orc.set_epa_bounds(abi.EPA_MAX_FACES, abi.EPA_MAX_HORIZON)
    st, _, _, _ = orc.gjk_epa(sc, hp, 64, 64, 1e-6)
    over = hp[st == 3]
    return sc, hp, st, over


def _child(sub_arrays, q):
    resource.setrlimit(resource.RLIMIT_AS, (MEMORY_LIMIT_B, MEMORY_LIMIT_B))
    sub = scenes.Scene(**sub_arrays)
    ref = O.CpuCollision("ref")
    t0 = time.perf_counter()
    st, nrm, dep, _ = ref.gjk_epa(sub, np.array([[0, 1]], np.uint32), 64, 64, 1e-6)
    q.put((int(st[0]), nrm[0].tolist(), float(dep[0]), time.perf_counter() - t0))


def run_reference(args):
    k, sub_arrays = args
    q = mp.get_context("fork").Queue()
    p = mp.get_context("fork").Process(target=_child, args=(sub_arrays, q))
    t0 = time.perf_counter()
    p.start()
    p.join(TIME_LIMIT_S)
    if p.is_alive():
        p.kill()
        p.join()
        return k, (-1, [0.0, 0.0, 0.0], 0.0, time.perf_counter() - t0)
    if q.empty():  # killed by the memory limit (std::bad_alloc ends the child)
        return k, (-1, [0.0, 0.0, 0.0], 0.0, time.perf_counter() - t0)
    return k, q.get()


def main():
    O.build(ref=True)
    sc, hp, st, over = overflow_pairs()
    print(f"{len(hp)} hull pairs; bounded oracle: {int((st == 0).sum())} misses, "
          f"{int((st == 1).sum())} converged, {int((st == 2).sum())} failures, "
          f"{len(over)} overflows", flush=True)
    subs = []
    for a, b in over:
        s = sc.subset(np.array([a, b]))
        subs.append(dict(pose=s.pose, shape_type=s.shape_type, shape_data=s.shape_data,
                         hull_range=s.hull_range, hull_points=s.hull_points))
    workers = int(os.environ.get("EPA_REF_WORKERS", "6"))
    results = {}
    if CACHE and os.path.exists(CACHE):
        with open(CACHE) as f:
            for line in f:
                e = json.loads(line)
                if tuple(e["pair"]) == tuple(int(x) for x in over[e["k"]]):
                    results[e["k"]] = tuple(e["result"])
    todo = [(k, s) for k, s in enumerate(subs) if k not in results]
    with cf.ThreadPoolExecutor(workers) as pool:  # each thread runs one child process
        for fut in cf.as_completed([pool.submit(run_reference, a) for a in todo]):
            k, r = fut.result()
            results[k] = r
            if CACHE:
                with open(CACHE, "a") as f:
                    f.write(json.dumps(dict(k=k, pair=[int(x) for x in over[k]],
                                            result=list(r))) + "\n")
            print(f"pair {k} {tuple(over[k])}: status {r[0]} depth {r[2]:.6g} "
                  f"({r[3]:.1f} s) [{len(results)}/{len(over)}]", flush=True)
    m = len(over)
    out = dict(pairs=over.astype(np.uint32),
               ref_status=np.array([results[k][0] for k in range(m)], np.int32),
               ref_normal=np.array([results[k][1] for k in range(m)], np.float32).reshape(m, 3),
               ref_depth=np.array([results[k][2] for k in range(m)], np.float32),
               ref_seconds=np.array([results[k][3] for k in range(m)], np.float64),
               time_limit_s=np.float64(TIME_LIMIT_S), memory_limit_b=np.int64(MEMORY_LIMIT_B),
               scene=np.array([N, SEED], np.int64), side=np.float64(SIDE),
               hull_pair_count=np.int64(len(hp)))
    # The two colliders of each pair, so CPU tests need not rebuild the 1M scene.
    sub = sc.subset(over.reshape(-1))
    out.update(pose=sub.pose, shape_type=sub.shape_type, shape_data=sub.shape_data,
               hull_range=sub.hull_range, hull_points=sub.hull_points)
    np.savez_compressed(os.path.join(HERE, "epa_overflow.npz"), **out)
    s = out["ref_status"]
    print(f"reference on {m} overflow pairs: {int((s == 1).sum())} converged, "
          f"{int((s == 2).sum())} failed, {int((s == 0).sum())} GJK misses, "
          f"{int((s == -1).sum())} not finished", flush=True)


if __name__ == "__main__":
    main()
