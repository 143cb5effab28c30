#!/usr/bin/env python3
"""Collision-step benchmark on one MI355X: 1,048,576 mixed bodies.

A step is the pose refresh (the integrated poses a stepping loop hands over,
read from device memory by physgpu_update_poses) plus one pass of the
collision section of PhysicsWorld::step over device-resident colliders: AABB
refresh, broad phase (SAP, uniform grid or LBVH), canonical pair order,
narrow phase (SS, SB, BB SAT + manifold, GJK/EPA for hulls) and ordered
manifold compaction.  Inputs are generated once (seeded) and uploaded before
timing.

Prints one JSON line (see the "contract" fields) to stdout; details go to
stderr.  `value` is the ms/step of the fastest broad phase (all three return
the same pair set); the per-algorithm numbers are in `per_algorithm`.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "phys-engine_amd"))

import numpy as np  # noqa: E402

import physgpu  # noqa: E402
from physgpu import abi, scenes  # noqa: E402

ALGOS = {"sap": abi.SWEEP_AND_PRUNE, "grid": abi.UNIFORM_GRID, "tree": abi.DYNAMIC_TREE}
HBM_PEAK_GBS = 8000.0
VALU_FP32_PEAK_TFLOPS = 157.3

# HBM bytes per dispatch measured with two rocprofv3 --pmc passes (FETCH_SIZE,
# WRITE_SIZE) over `bench.py --algorithms grid` on MI355X, scaled by the
# readings of a 256 MiB device copy (fetch x2.000, write x1.000); the summary
# with its commands is profiles/grid_step_mixed1M_hbm_traffic.md.  Valid for
# the 1,048,576-body mixed scene only.
MEASURED_TRAFFIC_1M = {
    "grid_query": 60.76e6,
    "grid_buckets": 185.60e6,
    "aabb_grid_keys": 160.54e6,
}


# Kernels of the broad phase (timer names of the library's launches).
BROADPHASE_PREFIXES = ("aabb_", "grid_", "radix_", "sap_", "lbvh_", "gather_")


def kernel_launches(per_algo, name):
    return per_algo.get("launches", {}).get(name, 1)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(name, n, pairs, hull_bodies, hull_points, table_entries):
    """Algorithmic (compulsory) HBM bytes of one launch, for the roofline.

    Per body: pose 28 B + type 1 B + shape 12 B read, AABB 24 B written
    (with the grid's keys: an 8-byte sort item and a 4-byte flag more);
    sorted records are 24 B of bounds + 4 B of index; sort items are 8 B;
    a pair record is 8 B; the grid's bucket pass reads an item and a 32-byte
    record and writes the record per body, and writes every 8-byte entry of
    the cell table.  The grid query of a collision step also classifies the
    pairs it emits: per pair a 4-byte hit flag and a 4-byte shape-class
    bucket entry written (the shape types ride in the records)."""
    return {
        "aabb_from_pose": n * 65 + hull_bodies * 8 + hull_points * 12,
        "aabb_grid_keys": n * (65 + 8 + 4) + hull_bodies * 8 + hull_points * 12,
        "aabb_reduce": n * 24,
        "sap_keys": n * (4 + 8),
        "gather_sorted": n * (8 + 24 + 28),
        "sap_sweep": n * 28 + pairs * 8,
        "grid_keys": n * (24 + 8 + 4),
        "grid_buckets": n * (8 + 32 + 32) + table_entries * 8,
        "grid_query": n * 28 + pairs * (8 + 4 + 4),
        "lbvh_morton": n * (24 + 8),
        "lbvh_query": n * 28 + pairs * 8,
    }.get(name)


def run_algorithm(pipe, algo, steps, warmup, pose_dev):
    """A step = the pose refresh from device memory (physgpu_update_poses,
    the integrated poses a stepping loop hands over) + the collision section.
    Eager launches, then the collision section replayed from a HIP graph (one
    launch per step), then per-kernel HIP-event timings of eager steps."""
    def step(graph):
        pipe.update_poses_device(pose_dev)
        if graph:
            pipe.collide_graph(algorithm=algo, sync=False)
        else:
            pipe.collide(algorithm=algo, sync=False)

    for _ in range(warmup):
        step(False)
    pipe.step_counts()
    t0 = time.perf_counter()
    for _ in range(steps):
        step(False)
    npairs, nman = pipe.step_counts()  # synchronises the stream
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / steps
    for _ in range(warmup):
        step(True)
    pipe.step_counts()
    t0 = time.perf_counter()
    for _ in range(steps):
        step(True)
    gpairs, gman = pipe.step_counts()
    graph_ms = (time.perf_counter() - t0) * 1e3 / steps
    if (gpairs, gman) != (npairs, nman):
        raise RuntimeError(f"graph replay counts {gpairs, gman} differ from eager {npairs, nman}")
    # Second pass with per-kernel HIP events on the library's stream.
    pipe.set_timing(True)
    for _ in range(steps):
        step(False)
    pipe.step_counts()
    kt = pipe.kernel_times()
    pipe.set_timing(False)
    return ms, graph_ms, npairs, nman, kt


def host_info():
    """CPU of the host the baseline runs on (nproc, affinity, model)."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "cpu": model}


def cpu_baseline(scene, steps=3):
    """The oracle's CPU restatement of the same collision step (bit-exact port
    of the reference's collision section): grid broad phase single-threaded as
    in the reference, the narrow phase on a worker pool of the host's share of
    cores (at most 16 per GPU); median of `steps` steps after one warm-up."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    orc = O.CpuCollision("oracle")
    # EPA as the GPU runs it (the reference's list within the LDS bounds,
    # continued past them over distinct faces with copy counts): the
    # reference's own list takes minutes on a few hull pairs.
    orc.mirror_gpu_epa()
    host = host_info()
    cores = max(1, min(host["affinity"], 16))
    w = scenes.with_world_extras(scenes.Scene(pose=scene.pose, shape_type=scene.shape_type,
                                              shape_data=scene.shape_data,
                                              hull_range=scene.hull_range,
                                              hull_points=scene.hull_points),
                                 local_rotation=True)
    t0 = time.perf_counter()
    m, t, s = orc.world_collide(w, algorithm=abi.UNIFORM_GRID, bp_workers=1, np_workers=cores,
                                warmup=1, steps=steps, capacity=4 * scene.count)
    wall = time.perf_counter() - t0
    step_ms = t["broadphase_ms"] + t["narrowphase_ms"] + t["aabb_ms"]
    return {"value": round(step_ms, 3), "unit": "ms/step", "cores": cores, "kind": "port",
            "host": host,
            "sample": f"median of {steps} steps (after 1 warm-up) of the full {scene.count}-body "
                      f"workload: AABB refresh + grid broad phase (1 thread, as the reference) + "
                      f"narrow phase ({cores} threads); broadphase {t['broadphase_ms']:.1f} ms, "
                      f"narrowphase {t['narrowphase_ms']:.1f} ms, AABB {t['aabb_ms']:.1f} ms; "
                      f"{len(m)} manifolds; wall {wall:.1f} s"}


def fp64_error(scene, pairs, man):
    """|FP32 - FP64| of the step's manifolds, per shape-pair class: the GPU's
    records against the oracle's FP64 re-evaluation of the same pairs (the
    same algorithm in double), over the manifolds where FP64 reaches the same
    outcome (a contact with the same point count).  Normals, depths and the
    collider-local contact anchors (A's and B's).  EPA against a sphere's
    curved support stops on whichever polytope face meets the 1e-6 test first,
    so FP64 can settle on another face there: the sphere-hull class reports
    the share of manifolds whose normals agree to 1e-4 besides its maxima."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    orc = O.CpuCollision("oracle")
    orc.mirror_gpu_epa()
    t0 = time.perf_counter()
    rec, hit = orc.generate_contacts_f64(scene, pairs)
    k = man["pair_index"].astype(np.int64)
    ok = hit[k] & (rec[k, 0] == man["point_count"])
    names = ("sphere_sphere", "sphere_box", "box_box", "sphere_hull", "box_hull", "hull_hull")
    out = {"manifolds": int(len(man)), "compared": int(ok.sum()),
           "note": "max |FP32 - FP64| over manifolds with the same FP64 outcome; anchors are "
                   "collider-local contact points (m)"}
    for cls, name in enumerate(names):
        sel = ok & (man["shape_pair"] == cls)
        if not sel.any():
            continue
        r, g = rec[k[sel]], man[sel]
        used = np.arange(4)[None, :] < g["point_count"][:, None]
        d64 = r[:, 4:32].reshape(-1, 4, 7)
        n_err = np.abs(r[:, 1:4] - g["normal"]).max(1)
        a_err = max(np.abs(d64[:, :, 0:3] - g["anchor_a"])[used].max(),
                    np.abs(d64[:, :, 3:6] - g["anchor_b"])[used].max())
        out[name] = {"compared": int(sel.sum()), "normal": float(n_err.max()),
                     "depth": float(np.abs(d64[:, :, 6] - g["penetration"])[used].max()),
                     "anchor": float(a_err),
                     "normal_within_1e-4": round(float((n_err <= 1e-4).mean()), 4)}
    log(f"fp64: {int(ok.sum())} of {len(man)} manifolds compared in "
        f"{time.perf_counter() - t0:.1f} s")
    return out


CLASS_NAMES = ("sphere_sphere", "sphere_box", "box_box", "sphere_hull", "box_hull", "hull_hull")


def flop_count(scene, pairs, cores):
    """FP32 operations of the step's narrow phase: the oracle's op-counting
    scalar (add/subtract, multiply, divide, square root of the reference
    algorithm; EPA as the GPU runs it past the LDS bounds) over every
    candidate pair, per shape-pair class, with GJK / EPA iteration counts."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    orc = O.CpuCollision("oracle")
    orc.mirror_gpu_epa()
    t0 = time.perf_counter()
    f = orc.count_flops(scene, pairs, workers=cores).astype(np.int64)
    t = scene.shape_type
    ta, tb = t[pairs[:, 0]].astype(np.int64), t[pairs[:, 1]].astype(np.int64)
    lo, hi = np.minimum(ta, tb), np.maximum(ta, tb)
    cls = np.where(hi < 2, lo + hi, 3 + lo)
    total = f[:, :4].sum(axis=1)
    out = {}
    for c, name in enumerate(CLASS_NAMES):
        m = cls == c
        if not m.any():
            continue
        e = {"pairs": int(m.sum()), "flops": int(total[m].sum()),
             "flops_per_pair": round(float(total[m].mean()), 1)}
        if c >= 3:
            gi, ei = int(f[m, 5].sum()), int(f[m, 6].sum())
            gf = int(f[m, 4].sum())
            e.update({"gjk_iterations": gi, "epa_iterations": ei,
                      "gjk_flops_per_iteration": round(gf / max(gi, 1), 1),
                      "epa_flops_per_iteration": round((int(total[m].sum()) - gf) / max(ei, 1), 1)})
        out[name] = e
    log(f"flop count: {len(pairs)} pairs in {time.perf_counter() - t0:.1f} s")
    return out


def roofline_narrowphase(per_class, kernels):
    """FP32 VALU roofline of the narrow phase: the dominant kernel (GJK + EPA
    of the hull pairs; its flops include the iterations continued past the LDS
    bounds by the consumer kernels) and the whole narrow phase (every class
    over the narrow-phase stage of the step, classification to compaction,
    whose SS / SB / BB kernels and canonical sort overlap GJK + EPA on other
    streams), against MI355X's FP32 vector peak."""
    hull = sum(v["flops"] for k, v in per_class.items() if k.endswith("hull"))
    allf = sum(v["flops"] for v in per_class.values())
    us_gjk = kernels.get("np_gjk_epa")
    us_np = kernels.get("stage_narrowphase", 0.0)
    peak = VALU_FP32_PEAK_TFLOPS

    def rate(fl, us):
        return round(fl / (us * 1e-6) / 1e12, 4) if us else None

    return {"bound": "valu", "kernel": "np_gjk_epa", "flops": hull, "avg_us": us_gjk,
            "achieved": rate(hull, us_gjk), "peak": peak, "unit": "TFLOP/s",
            "frac": round(rate(hull, us_gjk) / peak, 6) if us_gjk else None,
            "step": {"flops": allf, "avg_us": round(us_np, 2), "achieved": rate(allf, us_np),
                     "frac": round(rate(allf, us_np) / peak, 6) if us_np else None},
            "per_class": per_class,
            "note": "FP32 add/sub/mul/div/sqrt of the reference algorithm counted by the oracle's "
                    "op-counting scalar; no FMA (the reference's operation order) and no packed "
                    "FP32, so a quarter of the peak is this code's ceiling"}


def _hip():
    import ctypes as C
    return C.CDLL("libamdhip64.so")


def batched_bench(worlds, steps, warmup, check_every=64):
    """The batched-worlds form against per-world launches of the plain pipeline.

    Batch: `worlds` mixed worlds of log-uniform size in [16, 4096] bodies in one
    context, one launch per phase, step replayed from a HIP graph after an
    autotuning pass.  Per-world: one plain context per world (colliders
    resident), physgpu_collide called world after world on one stream.
    Both run the uniform-grid broad phase.  Returns a dict for the JSON line."""
    import ctypes as C
    t0 = time.perf_counter()
    sc, off, counts = scenes.batched_worlds(worlds)
    log(f"batched: {worlds} worlds, {sc.count} bodies (generated in "
        f"{time.perf_counter() - t0:.1f} s)")
    grid = abi.UNIFORM_GRID
    pipe = physgpu.CollisionPipeline(sc.count, max_pairs=4 * sc.count,
                                     max_hull_points=sc.hull_points.shape[1])
    pipe.set_colliders(sc)
    pipe.set_worlds(off)
    t0 = time.perf_counter()
    tuning = pipe.autotune(algorithm=grid, repeats=2)
    log(f"batched: autotune {time.perf_counter() - t0:.1f} s -> {tuning}")
    for _ in range(warmup):
        pipe.collide_graph(algorithm=grid, sync=False)
    pipe.step_counts()
    t0 = time.perf_counter()
    for _ in range(steps):
        pipe.collide_graph(algorithm=grid, sync=False)
    npairs, nman = pipe.step_counts()
    graph_ms = (time.perf_counter() - t0) * 1e3 / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        pipe.collide(algorithm=grid, sync=False)
    pipe.step_counts()
    eager_ms = (time.perf_counter() - t0) * 1e3 / steps
    pairs_b, man_b = pipe.pairs(), pipe.manifolds()
    po, mo = pipe.world_ranges()
    log(f"batched: graph {graph_ms:.3f} ms/step, eager {eager_ms:.3f} ms/step, "
        f"{npairs} pairs, {nman} manifolds")

    # Per-world contexts on one shared stream, for a sample of every
    # `sample_every`-th world (at most 1,024 contexts alive); the loop time is
    # scaled to all worlds by the sample's share of the batch's bodies.
    hip = _hip()
    stream = C.c_void_p()
    assert hip.hipStreamCreateWithFlags(C.byref(stream), C.c_uint(1)) == 0
    t0 = time.perf_counter()
    sample_every = max(1, -(-worlds // 1024))
    sample = list(range(0, worlds, sample_every))
    ctxs = []
    for w in sample:
        sub = sc.subset(np.arange(off[w], off[w + 1]))
        nw = sub.count
        p = physgpu.CollisionPipeline(nw, max_pairs=max(4 * nw, 1024),
                                      max_hull_points=max(sub.hull_points.shape[1], 1))
        p.set_stream(stream.value)
        p.set_colliders(sub)
        ctxs.append(p)
    log(f"per-world: {len(ctxs)} contexts set up in {time.perf_counter() - t0:.1f} s")
    for p in ctxs:
        p.collide(algorithm=grid, sync=False)
    assert hip.hipStreamSynchronize(stream) == 0
    loops = max(1, min(steps, 3))
    t0 = time.perf_counter()
    for _ in range(loops):
        for p in ctxs:
            p.collide(algorithm=grid, sync=False)
    assert hip.hipStreamSynchronize(stream) == 0
    sample_ms = (time.perf_counter() - t0) * 1e3 / loops
    sample_bodies = int(sum(counts[w] for w in sample))
    loop_ms = sample_ms * worlds / len(sample)
    # Identical results: counts of every sampled world, all words of every
    # check_every-th world.
    same = True
    for w, p in zip(sample, ctxs):
        pw, mw = p.step_counts()
        if pw != po[w + 1] - po[w] or mw != mo[w + 1] - mo[w]:
            same = False
        if w % check_every == 0:
            pp = p.pairs().astype(np.int64) + int(off[w])
            same &= bool(np.array_equal(pp, pairs_b[po[w]:po[w + 1]].astype(np.int64)))
            mm = p.manifolds().view(np.uint32).reshape(-1, 40)[:, 2:36]
            mb = man_b[mo[w]:mo[w + 1]].view(np.uint32).reshape(-1, 40)[:, 2:36]
            same &= bool(mm.shape == mb.shape and np.array_equal(mm, mb))
    for p in ctxs:
        p.close()
    hip.hipStreamDestroy(stream)
    pipe.close()
    log(f"per-world loop: {sample_ms:.1f} ms per pass over {len(sample)} sampled worlds "
        f"({sample_bodies} bodies), {loop_ms:.1f} ms scaled to {worlds}; identical: {same}")
    return {
        "workload": f"{worlds} worlds, bodies log-uniform in [16, 4096] ({sc.count} total), "
                    "mixed shapes, ~1 pair per body",
        "bodies": sc.count, "pairs": int(npairs), "manifolds": int(nman),
        "graph_ms_per_step": round(graph_ms, 4), "eager_ms_per_step": round(eager_ms, 4),
        "worlds_per_s": round(worlds / (graph_ms * 1e-3), 1),
        "per_world_loop_ms": round(loop_ms, 3),
        "per_world_sample": f"every {sample_every}th world ({len(sample)} contexts, "
                            f"{sample_bodies} bodies): {sample_ms:.3f} ms, scaled by "
                            f"{worlds}/{len(sample)} worlds",
        "per_world_worlds_per_s": round(worlds / (loop_ms * 1e-3), 1),
        "speedup_vs_per_world": round(loop_ms / graph_ms, 2),
        "identical_to_per_world": same,
        "tuning": tuning,
    }


def adjoint_bench(pairs, steps):
    """Narrow-phase adjoint on `pairs` isolated touching pairs (half box-box,
    half sphere-box, depth U[1e-3, 5e-2] m): loss = sum depth^2 + sum normal . w.
    Device time of the adjoint kernel from HIP events (upstream gradients are
    staged to the device before it, outside the kernel time)."""
    sc = scenes.contact_pairs(pairs)
    pipe = physgpu.CollisionPipeline(sc.count, 4 * sc.count)
    pipe.set_colliders(sc)
    _, m = pipe.collide(algorithm=abi.UNIFORM_GRID)
    man = pipe.manifolds()
    rng = np.random.default_rng(1)
    w = rng.standard_normal((3, m))
    w = (w / np.linalg.norm(w, axis=0)).astype(np.float32)
    used = np.arange(4)[:, None] < man["point_count"][None, :]
    gd = np.where(used, 2.0 * man["penetration"].T, 0.0).astype(np.float32)
    pipe.contact_adjoint(w, gd)
    pipe.set_timing(True)
    for _ in range(steps):
        pipe.contact_adjoint(w, gd)
    kt = pipe.kernel_times()
    pipe.collide(algorithm=abi.UNIFORM_GRID)
    kf = pipe.kernel_times()
    pipe.set_timing(False)
    pipe.close()
    adj_us = kt["np_adjoint"][0] * 1e3
    fwd_us = sum(kf[k][0] for k in ("np_sphere_box", "np_box_box") if k in kf) * 1e3
    log(f"adjoint: {m} manifolds, kernel {adj_us:.1f} us (forward SB+BB kernels {fwd_us:.1f} us)")
    return {"workload": f"{m} touching pairs, half box-box, half sphere-box",
            "manifolds": int(m), "adjoint_kernel_us": round(adj_us, 2),
            "forward_kernels_us": round(fwd_us, 2),
            "manifolds_per_s": round(m / (adj_us * 1e-6), 1)}


def tree_bench(bounds, steps):
    """Persistent proxies (physgpu_tree_*) over the step's AABBs: create every
    proxy, then per step move every proxy (x + 0.01 m), update all of them
    and query, with the LBVH rebuilt or refitted.  Host-API wall time of
    synchronous calls (bounds staged from host memory, pairs copied back)."""
    n = len(bounds)
    t = physgpu.DynamicAabbTree(capacity=n, max_pairs=max(4 * n // 8, 1 << 16))
    users = np.arange(n, dtype=np.uint64)
    t0 = time.perf_counter()
    ids = t.create_proxies(bounds, users)
    create_ms = (time.perf_counter() - t0) * 1e3
    out = {}
    b = bounds.copy()
    pairs = None
    for policy in ("rebuild", "refit"):
        t.policy = abi.TREE_REFIT if policy == "refit" else abi.TREE_REBUILD
        t.find_candidate_pairs()
        upd = qry = 0.0
        for _ in range(steps):
            b[:, 0] += np.float32(0.01)
            b[:, 3] += np.float32(0.01)
            t0 = time.perf_counter()
            t.update_proxies(ids, b)
            t1 = time.perf_counter()
            got = t.find_candidate_pairs()
            t2 = time.perf_counter()
            upd += t1 - t0
            qry += t2 - t1
        pairs = len(got)
        out[policy] = {"update_ms": round(upd * 1e3 / steps, 3),
                       "query_ms": round(qry * 1e3 / steps, 3)}
    t.close()
    log(f"proxy tree: {n} proxies, {pairs} pairs; create {create_ms:.1f} ms; "
        f"rebuild query {out['rebuild']['query_ms']:.2f} ms, refit query "
        f"{out['refit']['query_ms']:.2f} ms, update {out['refit']['update_ms']:.2f} ms")
    return {"workload": f"{n} proxies on the step's AABBs, all moved every step",
            "proxies": n, "pairs": pairs, "create_ms": round(create_ms, 3), **out}


def copy_probe(nbytes=256 << 20):
    """One device-to-device copy of known size (the HIP runtime's blit kernel),
    used to scale rocprof FETCH_SIZE / WRITE_SIZE readings."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy probe: {nbytes} bytes read + {nbytes} bytes written")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--bodies", type=int, default=1048576)
    ap.add_argument("--algorithms", default="sap,grid,tree")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batched-worlds", type=int, default=4096,
                    help="worlds of the batched-form comparison (0: skip it)")
    ap.add_argument("--adjoint-pairs", type=int, default=65536,
                    help="touching pairs of the adjoint timing (0: skip it)")
    ap.add_argument("--tree-proxies", type=int, default=1048576,
                    help="proxies of the persistent-tree timing (0: skip it)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one 256 MiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        return 2
    if args.copy_probe:
        copy_probe()
    n = args.bodies
    side = 256.0 * (n / 1048576.0) ** (1.0 / 3.0)
    log(f"scene: {n} mixed bodies, side {side:.1f} m")
    scene = scenes.mixed(n, seed=20261203, side=side)
    hull_bodies = int((scene.shape_type == abi.HULL).sum())
    hull_points = 0 if scene.hull_points is None else scene.hull_points.shape[1]
    pipe = physgpu.CollisionPipeline(n, max_pairs=max(4 * n // 8, 1 << 16),
                                     max_hull_points=max(hull_points, 1))
    pipe.set_colliders(scene)
    # The poses a stepping loop hands over every step, resident on the device.
    import ctypes as C
    hip = _hip()
    pose_dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(pose_dev), C.c_size_t(scene.pose.nbytes)) == 0
    assert hip.hipMemcpy(pose_dev, scene.pose.ctypes.data_as(C.c_void_p),
                         C.c_size_t(scene.pose.nbytes), 1) == 0  # host to device
    # Grid cell table: 2^3k entries, the smallest k >= 2 with 2^3k >= 2 n.
    k = 2
    while k < 9 and (1 << (3 * k)) < 2 * n:
        k += 1
    table_entries = 1 << (3 * k)

    per = {}
    for name in args.algorithms.split(","):
        algo = ALGOS[name]
        eager_ms, graph_ms, npairs, nman, kt = run_algorithm(pipe, algo, args.steps, args.warmup,
                                                             pose_dev.value)
        ms = min(eager_ms, graph_ms)
        per[name] = {"ms_per_step": round(ms, 4), "eager_ms_per_step": round(eager_ms, 4),
                     "graph_ms_per_step": round(graph_ms, 4), "pairs": npairs, "manifolds": nman,
                     "kernels": {k: round(v[0] * 1e3, 2) for k, v in kt.items()},
                     "launches": {k: v[1] // args.steps for k, v in kt.items()}}
        log(f"{name}: {ms:.3f} ms/step (eager {eager_ms:.3f}, graph replay {graph_ms:.3f}), "
            f"{npairs} pairs, {nman} manifolds")
        for kn, (avg, launches) in sorted(kt.items(), key=lambda x: -x[1][0] * x[1][1]):
            log(f"   {kn:24s} {avg * 1e3:9.2f} us x {launches // args.steps}")
    best = min(per, key=lambda k: per[k]["ms_per_step"])
    ms = per[best]["ms_per_step"]
    step_pairs, step_man = pipe.pairs(), pipe.manifolds()  # the same for every broad phase
    npairs = per[best]["pairs"]

    # Roofline of the dominant bandwidth kernel of the headline step.
    kt = per[best]["kernels"]
    ranked = sorted(kt.items(), key=lambda x: -x[1])
    roof = None
    for kn, us in ranked:
        b = kernel_bytes(kn, n, npairs, hull_bodies, hull_points, table_entries)
        if b is None:
            continue
        achieved = b / (us * 1e-6) / 1e9
        traffic = MEASURED_TRAFFIC_1M.get(kn) if n == 1048576 else None
        roof = {"bound": "hbm", "kernel": kn, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "bytes_per_launch": b, "avg_us": us}
        break
    # The broad phase as a whole against its compulsory bytes B_bp = 108 N + 8 P
    # (pose + shape read 44 B, AABB write 24 B, sorted AABB read 24 B, one
    # key/index sort round trip 16 B per body; 8 B per pair written), over the
    # summed device time of its kernels (AABB refresh through pair emission;
    # the canonical pair order runs on the side stream during the narrow phase).
    # The grid query classifies its pairs as well: 8 B more per pair.
    bp_names = [kn for kn in kt if kn.startswith(BROADPHASE_PREFIXES)]
    bp_us = sum(kt[kn] * kernel_launches(per[best], kn) for kn in bp_names)
    bp_bytes = 108 * n + (16 if best == "grid" else 8) * npairs
    roof_bp = {"bound": "hbm", "kernels": sorted(bp_names), "avg_us": round(bp_us, 2),
               "bytes": bp_bytes,
               "achieved": round(bp_bytes / (bp_us * 1e-6) / 1e9, 1) if bp_us else None,
               "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": round(bp_bytes / (bp_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4) if bp_us else None}

    line = {
        "metric": "collision ms/step at 1M bodies",
        "value": ms,
        "unit": "ms/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"mixed-{n}: 50% spheres r U[0.2,0.5], 40% boxes h U[0.2,0.6], "
                               f"10% 16-point hulls, random orientations, {side:.0f} m cube",
                   "bodies": n, "broadphase": best, "candidate_pairs": npairs,
                   "manifolds": per[best]["manifolds"], "parallelism": "1 GPU",
                   "execution": ("HIP graph replay of the full step"
                                 if per[best]["graph_ms_per_step"] <= per[best]["eager_ms_per_step"]
                                 else "eager launches")},
        "pairs_per_s": round(npairs / (ms * 1e-3), 1),
        "manifolds_per_s": round(per[best]["manifolds"] / (ms * 1e-3), 1),
        "per_algorithm": {k: {kk: vv for kk, vv in v.items() if kk not in ("kernels", "launches")}
                          for k, v in per.items()},
        "roofline": roof,
        "roofline_broadphase": roof_bp,
        "cpu_baseline": None,
    }
    if args.batched_worlds > 0:
        line["batched"] = batched_bench(args.batched_worlds, args.steps, args.warmup)
    if args.adjoint_pairs > 0:
        line["adjoint"] = adjoint_bench(args.adjoint_pairs, args.steps)
    if args.tree_proxies > 0:
        aabbs = pipe.compute_aabbs()[:args.tree_proxies]
        line["proxy_tree"] = tree_bench(np.ascontiguousarray(aabbs, np.float32), min(args.steps, 5))
    if not args.no_cpu_baseline:
        # The CPU leg (the oracle, as checker and baseline only): the CPU step
        # time, the error against an FP64 re-evaluation of the headline
        # step's manifolds, and the operation count of its narrow phase.
        try:
            line["cpu_baseline"] = cpu_baseline(scene)
        except Exception as e:  # reported, never fatal for the GPU number
            line["cpu_baseline"] = {"error": repr(e)}
        try:
            line["fp64_max_error"] = fp64_error(scene, step_pairs, step_man)
        except Exception as e:
            line["fp64_max_error"] = {"error": repr(e)}
        try:
            cores = max(1, min(host_info()["affinity"], 16))
            line["roofline_narrowphase"] = roofline_narrowphase(
                flop_count(scene, step_pairs, cores), per[best]["kernels"])
        except Exception as e:
            line["roofline_narrowphase"] = {"error": repr(e)}
    log(json.dumps(per, indent=1))
    print(json.dumps(line), flush=True)
    pipe.close()
    hip.hipFree(pose_dev)
    return 0


if __name__ == "__main__":
    sys.exit(main())
