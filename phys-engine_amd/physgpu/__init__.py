"""Host-side mirror of phys-engine's collision interfaces over libphysgpu.so.

Names and argument meaning follow the reference:
  BroadPhaseAlgorithm                 include/phys/collision/broadphase.hpp:8-13
  BroadPhase.find_candidate_pairs     broadphase.hpp:52-54 (raises ValueError for
                                      DynamicTree, as the reference throws
                                      std::invalid_argument, broadphase.cpp:200-201)
  NarrowPhase.generate_contacts       narrowphase.hpp:36-39, batched over pairs
  CollisionPipeline.collide           the collision section of PhysicsWorld::step
                                      (physics_world.cpp:539-739), device resident
All work runs in the HIP library; when it cannot be loaded every call raises.
"""
import ctypes as C
import enum
import os
import warnings

import numpy as np

from . import abi
from .abi import MANIFOLD_DTYPE

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libphysgpu.so")
_lib = None


class PhysGpuError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"physgpu status {status}: {message}")
        self.status = status


class CapacityError(PhysGpuError):
    pass


class BroadPhaseAlgorithm(enum.IntEnum):
    SweepAndPrune = abi.SWEEP_AND_PRUNE
    UniformGrid = abi.UNIFORM_GRID
    DynamicTree = abi.DYNAMIC_TREE


def library():
    """Load libphysgpu.so (built by phys-engine_amd/build.py); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with phys-engine_amd/build.py")
    L = C.CDLL(LIB_PATH)
    vp, u32, u64, i32, dbl = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_double
    L.physgpu_last_error.restype = C.c_char_p
    L.physgpu_device_count.restype = C.c_int
    sigs = {
        "physgpu_create": [C.POINTER(vp), i32, u32, u64, u32],
        "physgpu_destroy": [vp],
        "physgpu_set_stream": [vp, vp],
        "physgpu_set_timing": [vp, i32],
        "physgpu_set_colliders": [vp, C.POINTER(abi.Colliders), i32],
        "physgpu_compute_aabbs": [vp, vp],
        "physgpu_find_candidate_pairs": [vp, vp, u32, i32, dbl, vp, u64, C.POINTER(u64),
                                         C.POINTER(abi.BroadPhaseStats)],
        "physgpu_generate_contacts": [vp, vp, u64, C.POINTER(abi.NarrowPhaseOptions), vp, u64,
                                      C.POINTER(u64), C.POINTER(abi.NarrowPhaseStats)],
        "physgpu_collide": [vp, i32, dbl, C.POINTER(abi.NarrowPhaseOptions), C.POINTER(u64),
                            C.POINTER(u64)],
        "physgpu_set_world": [vp, vp, C.POINTER(abi.WorldExtras), i32],
        "physgpu_download_pairs": [vp, vp, u64],
        "physgpu_download_manifolds": [vp, vp, u64],
        "physgpu_narrowphase_stats_get": [vp, C.POINTER(abi.NarrowPhaseStats)],
        "physgpu_step_counts": [vp, C.POINTER(u64), C.POINTER(u64)],
        "physgpu_epa_unresolved": [vp, vp, u64, C.POINTER(u64)],
        "physgpu_step_stats_get": [vp, C.POINTER(abi.StepStats)],
        "physgpu_kernel_times": [vp, C.POINTER(C.c_char_p), C.POINTER(dbl), C.POINTER(u64), u32,
                                 C.POINTER(u32)],
        "physgpu_set_worlds": [vp, u32, vp, i32],
        "physgpu_update_poses": [vp, vp, i32],
        "physgpu_contact_adjoint": [vp, u64, vp, vp, vp, vp, vp, i32, C.POINTER(u64)],
        "physgpu_contact_geometry": [vp, u64, vp, vp, vp, i32],
        "physgpu_world_ranges": [vp, vp, vp],
        "physgpu_collide_graph": [vp, i32, dbl, C.POINTER(abi.NarrowPhaseOptions),
                                  C.POINTER(u64), C.POINTER(u64)],
        "physgpu_set_tuning": [vp, C.POINTER(abi.Tuning)],
        "physgpu_get_tuning": [vp, C.POINTER(abi.Tuning)],
        "physgpu_autotune": [vp, i32, u32, C.POINTER(abi.Tuning)],
        "physgpu_tree_create_proxies": [vp, u32, vp, vp, vp],
        "physgpu_tree_destroy_proxies": [vp, u32, vp],
        "physgpu_tree_update_proxies": [vp, u32, vp, vp, vp],
        "physgpu_tree_size": [vp, C.POINTER(u32)],
        "physgpu_tree_find_pairs": [vp, i32, vp, u64, C.POINTER(u64),
                                    C.POINTER(abi.BroadPhaseStats)],
    }
    for name, args in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _lib = L
    return L


class EpaUnresolvedWarning(RuntimeWarning):
    """Some hull pairs of a step got no EPA result (PHYSGPU_EPA_UNRESOLVED);
    CollisionPipeline.unresolved_pairs() lists them."""


class AdjointSkippedWarning(RuntimeWarning):
    """Some manifolds got zero gradients from physgpu_contact_adjoint:
    sphere-hull manifolds (their EPA polytope's vertices depend on earlier
    faces' directions, which the adjoint does not follow) and hull manifolds
    whose EPA ran past the LDS polytope bounds."""


def _warn_skipped(skipped, m):
    if skipped:
        warnings.warn(f"{skipped} of {m} manifolds have zero gradients (sphere-hull, or EPA "
                      f"past the LDS bounds)", AdjointSkippedWarning, stacklevel=3)


def _check(status, allow=()):
    if status == abi.OK or status in allow:
        return status
    msg = library().physgpu_last_error().decode(errors="replace")
    if status == abi.EPA_UNRESOLVED:  # every result is written
        warnings.warn(msg, EpaUnresolvedWarning, stacklevel=3)
        return status
    if status == abi.CAPACITY_EXCEEDED:
        raise CapacityError(status, msg)
    if status == abi.INVALID_ARGUMENT:
        raise ValueError(f"physgpu: {msg}")
    raise PhysGpuError(status, msg)


def device_count():
    return library().physgpu_device_count()


def _options(gjk_iterations=64, epa_iterations=64, epa_tolerance=1e-6):
    o = abi.NarrowPhaseOptions()
    o.gjk_max_iterations = gjk_iterations
    o.epa_max_iterations = epa_iterations
    o.epa_tolerance = epa_tolerance
    return o


def _pairs_array(raw):
    return raw.view(np.uint32).reshape(-1, 2)


class CollisionPipeline:
    """One device context: resident colliders, AABBs, pairs and manifolds."""

    def __init__(self, max_colliders, max_pairs=None, max_hull_points=0, device=0):
        L = library()
        self._ctx = C.c_void_p()
        max_pairs = max_pairs or max(4 * max_colliders, 1024)
        _check(L.physgpu_create(C.byref(self._ctx), device, max_colliders, max_pairs,
                                max_hull_points))
        self.max_colliders = max_colliders
        self.max_pairs = max_pairs
        self.count = 0
        self._keep = None

    def close(self):
        if self._ctx:
            library().physgpu_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._ctx

    def set_stream(self, hip_stream_handle):
        _check(library().physgpu_set_stream(self._ctx, C.c_void_p(hip_stream_handle)))

    def set_timing(self, enabled=True):
        _check(library().physgpu_set_timing(self._ctx, int(bool(enabled))))

    def set_colliders(self, scene):
        c = abi.colliders_struct(scene)
        self._keep = scene
        _check(library().physgpu_set_colliders(self._ctx, C.byref(c), abi.HOST))
        self.count = scene.count
        self.worlds = 0

    def update_poses(self, pose):
        """New poses (float32 [7, count]) for the resident colliders; shapes,
        worlds and a captured step graph are kept."""
        pose = np.ascontiguousarray(pose, np.float32)
        if pose.shape != (7, self.count):
            raise ValueError(f"pose must be [7, {self.count}]")
        _check(library().physgpu_update_poses(self._ctx, pose.ctypes.data, abi.HOST))

    def set_world(self, scene):
        """World-level form: body poses, body indices, local poses, materials."""
        n = scene.count
        body_pose = np.zeros((7, n), np.float32)
        body_pose[:3] = scene.body_position
        body_pose[6] = 1.0
        ex = abi.WorldExtras()
        body = getattr(scene, "body", None)
        if body is None:
            body = np.arange(n, dtype=np.uint32)
        keep = [body_pose, np.ascontiguousarray(body, np.uint32)]
        ex.body = abi.u32ptr(keep[1])
        ex.local_pose = abi.fptr(scene.local_pose)
        ex.friction = abi.fptr(scene.friction)
        ex.restitution = abi.fptr(scene.restitution)
        _check(library().physgpu_set_world(self._ctx, body_pose.ctypes.data, C.byref(ex),
                                           abi.HOST))
        self._world_keep = keep

    def compute_aabbs(self):
        out = np.zeros((max(self.count, 1), 6), np.float32)
        _check(library().physgpu_compute_aabbs(self._ctx, out.ctypes.data))
        return out[:self.count]

    def find_candidate_pairs(self, bounds=None, algorithm=BroadPhaseAlgorithm.SweepAndPrune,
                             cell_size=0.0, with_stats=False):
        L = library()
        count = C.c_uint64(0)
        stats = abi.BroadPhaseStats()
        if bounds is not None:
            bounds = np.ascontiguousarray(bounds, np.float32).reshape(-1, 6)
            n = bounds.shape[0]
            bptr = bounds.ctypes.data
        else:
            n, bptr = self.count, None
        out = np.zeros(self.max_pairs, np.uint64)
        _check(L.physgpu_find_candidate_pairs(self._ctx, bptr, n, int(algorithm), cell_size,
                                              out.ctypes.data, self.max_pairs, C.byref(count),
                                              C.byref(stats) if with_stats else None))
        if bounds is not None:
            self.count = n
        pairs = _pairs_array(out[:count.value])
        return (pairs, stats.as_dict()) if with_stats else pairs

    def generate_contacts(self, pairs=None, gjk_iterations=64, epa_iterations=64,
                          epa_tolerance=1e-6, with_stats=False):
        L = library()
        count = C.c_uint64(0)
        stats = abi.NarrowPhaseStats()
        opt = _options(gjk_iterations, epa_iterations, epa_tolerance)
        if pairs is not None:
            pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
            pptr, npairs = pairs.ctypes.data, pairs.shape[0]
            cap = max(npairs, 1)
        else:
            pptr, npairs, cap = None, 0, self.max_pairs
        out = np.zeros(cap, MANIFOLD_DTYPE)
        _check(L.physgpu_generate_contacts(self._ctx, pptr, npairs, C.byref(opt), out.ctypes.data,
                                           cap, C.byref(count), C.byref(stats)))
        m = out[:count.value]
        return (m, stats.as_dict()) if with_stats else m

    def collide(self, algorithm=BroadPhaseAlgorithm.SweepAndPrune, cell_size=0.0,
                gjk_iterations=64, epa_iterations=64, epa_tolerance=1e-6, sync=True):
        """Collision step on the resident colliders; returns (pairs, manifolds)
        counts when sync, else None (work stays queued on the stream)."""
        opt = _options(gjk_iterations, epa_iterations, epa_tolerance)
        if not sync:
            _check(library().physgpu_collide(self._ctx, int(algorithm), cell_size, C.byref(opt),
                                             None, None))
            return None
        p, m = C.c_uint64(0), C.c_uint64(0)
        _check(library().physgpu_collide(self._ctx, int(algorithm), cell_size, C.byref(opt),
                                         C.byref(p), C.byref(m)))
        return p.value, m.value

    def collide_graph(self, algorithm=BroadPhaseAlgorithm.UniformGrid, cell_size=0.0,
                      gjk_iterations=64, epa_iterations=64, epa_tolerance=1e-6, sync=True):
        """collide() replayed from a HIP graph (captured on the first call and
        again whenever the step's configuration changes)."""
        opt = _options(gjk_iterations, epa_iterations, epa_tolerance)
        L = library()
        if not sync:
            _check(L.physgpu_collide_graph(self._ctx, int(algorithm), cell_size, C.byref(opt),
                                           None, None))
            return None
        p, m = C.c_uint64(0), C.c_uint64(0)
        _check(L.physgpu_collide_graph(self._ctx, int(algorithm), cell_size, C.byref(opt),
                                       C.byref(p), C.byref(m)))
        return p.value, m.value

    def set_worlds(self, offsets):
        """Batch of worlds over the resident colliders: world w owns colliders
        [offsets[w], offsets[w + 1]).  None or [] returns to one world."""
        if offsets is None or len(offsets) == 0:
            _check(library().physgpu_set_worlds(self._ctx, 0, None, abi.HOST))
            self.worlds = 0
            return
        off = np.ascontiguousarray(offsets, np.uint32)
        _check(library().physgpu_set_worlds(self._ctx, len(off) - 1, off.ctypes.data, abi.HOST))
        self.worlds = len(off) - 1

    def world_ranges(self):
        """(pair_offsets, manifold_offsets) of the last step, uint64 [worlds + 1] each."""
        w = getattr(self, "worlds", 0)
        po = np.zeros(w + 1, np.uint64)
        mo = np.zeros(w + 1, np.uint64)
        _check(library().physgpu_world_ranges(self._ctx, po.ctypes.data, mo.ctypes.data))
        return po, mo

    def tuning(self):
        t = abi.Tuning()
        _check(library().physgpu_get_tuning(self._ctx, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, tuning=None):
        """Set launch shapes from a dict (missing or zero fields: defaults)."""
        if tuning is None:
            _check(library().physgpu_set_tuning(self._ctx, None))
            return
        t = abi.Tuning()
        for k, v in tuning.items():
            if isinstance(v, (list, tuple)):
                arr = getattr(t, k)
                for i, x in enumerate(v):
                    arr[i] = x
            else:
                setattr(t, k, v)
        _check(library().physgpu_set_tuning(self._ctx, C.byref(t)))

    def autotune(self, algorithm=BroadPhaseAlgorithm.UniformGrid, repeats=3):
        t = abi.Tuning()
        _check(library().physgpu_autotune(self._ctx, int(algorithm), repeats, C.byref(t)))
        return t.as_dict()

    def contact_geometry(self):
        """Differentiable outputs of the last step's manifolds, recomputed by
        the adjoint's forward: normal [3, m], depth [4, m], world points [12, m]."""
        _, m = self.step_counts()
        normal = np.zeros((3, max(m, 1)), np.float32)
        depth = np.zeros((4, max(m, 1)), np.float32)
        point = np.zeros((12, max(m, 1)), np.float32)
        _check(library().physgpu_contact_geometry(self._ctx, m, normal.ctypes.data,
                                                  depth.ctypes.data, point.ctypes.data, abi.HOST))
        return normal[:, :m], depth[:, :m], point[:, :m]

    def contact_adjoint(self, g_normal=None, g_depth=None, g_point=None, per_pair=False):
        """Vector-Jacobian product of the last step's manifolds: upstream
        gradients (planes [3, m], [4, m], [12, m]; None = 0) to pose gradients
        [7, colliders] (px py pz qx qy qz qw), or per manifold [14, m] when
        per_pair.  Returns (gradient, skipped hull manifolds)."""
        _, m = self.step_counts()

        def plane(a, rows):
            if a is None:
                return None
            a = np.ascontiguousarray(a, np.float32)
            if a.shape != (rows, m):
                raise ValueError(f"upstream gradient must be [{rows}, {m}]")
            return a
        gn, gd, gp = plane(g_normal, 3), plane(g_depth, 4), plane(g_point, 12)
        ptr = (lambda a: None if a is None else a.ctypes.data)
        skipped = C.c_uint64(0)
        if per_pair:
            out = np.zeros((14, max(m, 1)), np.float32)
            _check(library().physgpu_contact_adjoint(self._ctx, m, ptr(gn), ptr(gd), ptr(gp), None,
                                                     out.ctypes.data, abi.HOST,
                                                     C.byref(skipped)))
            _warn_skipped(skipped.value, m)
            return out[:, :m], skipped.value
        out = np.zeros((7, max(self.count, 1)), np.float32)
        _check(library().physgpu_contact_adjoint(self._ctx, m, ptr(gn), ptr(gd), ptr(gp),
                                                 out.ctypes.data, None, abi.HOST,
                                                 C.byref(skipped)))
        _warn_skipped(skipped.value, m)
        return out[:, :self.count], skipped.value

    # Device-pointer forms (torch tensors' data_ptr()); the caller orders its
    # own stream against these calls, which synchronise the context's stream.
    def update_poses_device(self, pose_ptr):
        _check(library().physgpu_update_poses(self._ctx, C.c_void_p(pose_ptr), abi.DEVICE))

    def contact_geometry_device(self, m, normal_ptr, depth_ptr, point_ptr):
        _check(library().physgpu_contact_geometry(self._ctx, m, C.c_void_p(normal_ptr),
                                                  C.c_void_p(depth_ptr), C.c_void_p(point_ptr),
                                                  abi.DEVICE))

    def contact_adjoint_device(self, m, g_normal_ptr, g_depth_ptr, g_point_ptr, grad_pose_ptr):
        skipped = C.c_uint64(0)
        _check(library().physgpu_contact_adjoint(
            self._ctx, m, C.c_void_p(g_normal_ptr), C.c_void_p(g_depth_ptr),
            C.c_void_p(g_point_ptr), C.c_void_p(grad_pose_ptr), None, abi.DEVICE,
            C.byref(skipped)))
        _warn_skipped(skipped.value, m)
        return skipped.value

    def step_counts(self):
        """(candidate pairs, manifolds) of the last step."""
        p, m = C.c_uint64(0), C.c_uint64(0)
        _check(library().physgpu_step_counts(self._ctx, C.byref(p), C.byref(m)))
        return p.value, m.value

    def pairs(self):
        """Canonical candidate pairs of the last broad phase, uint32 [P, 2]."""
        p, _ = self.step_counts()
        out = np.zeros(max(p, 1), np.uint64)
        _check(library().physgpu_download_pairs(self._ctx, out.ctypes.data, p))
        return _pairs_array(out[:p])

    def manifolds(self):
        """Manifolds of the last narrow phase, in candidate order."""
        _, m = self.step_counts()
        out = np.zeros(max(m, 1), MANIFOLD_DTYPE)
        _check(library().physgpu_download_manifolds(self._ctx, out.ctypes.data, m))
        return out[:m]

    def unresolved_pairs(self):
        """Hull pairs of the last narrow phase without an EPA result, uint32 [U, 2]."""
        L = library()
        count = C.c_uint64(0)
        _check(L.physgpu_epa_unresolved(self._ctx, None, 0, C.byref(count)))
        out = np.zeros((max(count.value, 1), 2), np.uint32)
        _check(L.physgpu_epa_unresolved(self._ctx, out.ctypes.data, count.value, C.byref(count)))
        return out[:count.value]

    def step_stats(self):
        """StepStats collision fields of the last step (dict)."""
        stats = abi.StepStats()
        _check(library().physgpu_step_stats_get(self._ctx, C.byref(stats)))
        return stats.as_dict()

    def narrowphase_stats(self):
        stats = abi.NarrowPhaseStats()
        _check(library().physgpu_narrowphase_stats_get(self._ctx, C.byref(stats)))
        return stats.as_dict()

    def kernel_times(self):
        L = library()
        cap = 64
        names = (C.c_char_p * cap)()
        avg = (C.c_double * cap)()
        launches = (C.c_uint64 * cap)()
        count = C.c_uint32(0)
        _check(L.physgpu_kernel_times(self._ctx, names, avg, launches, cap, C.byref(count)))
        return {names[i].decode(): (avg[i], launches[i]) for i in range(min(count.value, cap))}


class BroadPhase:
    """BroadPhase::findCandidatePairs (broadphase.hpp:52-54) on the GPU."""

    _pipe = None

    @classmethod
    def _pipeline(cls, n):
        if cls._pipe is None or cls._pipe.max_colliders < n:
            if cls._pipe is not None:
                cls._pipe.close()
            cls._pipe = CollisionPipeline(max(n, 1024), max(16 * n, 1 << 16))
        return cls._pipe

    @classmethod
    def find_candidate_pairs(cls, bounds, stats=None,
                             algorithm=BroadPhaseAlgorithm.SweepAndPrune):
        """bounds: float32 [N, 6] (min xyz, max xyz).  Returns uint32 [P, 2]
        lexicographic pairs; fills `stats` (a dict) when given."""
        if algorithm == BroadPhaseAlgorithm.DynamicTree:
            raise ValueError("DynamicTree requires a persistent DynamicAabbTree instance")
        bounds = np.ascontiguousarray(bounds, np.float32).reshape(-1, 6)
        pipe = cls._pipeline(bounds.shape[0])
        pairs, st = pipe.find_candidate_pairs(bounds, algorithm, with_stats=True)
        if stats is not None:
            stats.clear()
            stats.update(st)
        return pairs


class DynamicAabbTree:
    """DynamicAabbTree (dynamic_aabb_tree.hpp:25-32) with its proxies on the
    device (physgpu_tree_*): the reference's fat-bounds rule per update
    (dynamic_aabb_tree.cpp:44-57, 128-139), pairs of live proxies whose tight
    bounds overlap by user index (:268-336), from an LBVH rebuilt per query or
    refitted while the proxy set is unchanged (`policy`).  Proxy ids are
    recycled; invalid ids and bounds raise ValueError (the reference's
    std::invalid_argument)."""

    def __init__(self, capacity=1 << 16, max_pairs=1 << 20, policy="rebuild"):
        if policy not in ("rebuild", "refit"):
            raise ValueError("policy is 'rebuild' or 'refit'")
        self._pipe = CollisionPipeline(capacity, max_pairs)
        self.capacity = capacity
        self.policy = abi.TREE_REFIT if policy == "refit" else abi.TREE_REBUILD

    def close(self):
        if self._pipe is not None:
            self._pipe.close()
            self._pipe = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _aabbs(bounds):
        b = np.ascontiguousarray(np.asarray(bounds, np.float32).reshape(-1, 6))
        return b, b.ctypes.data

    def create_proxies(self, bounds, user_index):
        """createProxy for each row of bounds (min xyz, max xyz); proxy ids."""
        b, bp = self._aabbs(bounds)
        users = np.ascontiguousarray(np.asarray(user_index, np.uint64).reshape(-1))
        if len(users) != len(b):
            raise ValueError("one user index per bounds row")
        ids = np.zeros(len(b), np.uint32)
        _check(library().physgpu_tree_create_proxies(self._pipe._ctx, len(b), bp,
                                                      users.ctypes.data, ids.ctypes.data))
        return ids

    def create_proxy(self, bounds, user_index):
        return int(self.create_proxies([bounds], [user_index])[0])

    @staticmethod
    def _ids(proxies):
        """Proxy ids as uint32; wider values are invalid ids, never truncated
        onto live ones."""
        raw = np.asarray(proxies).reshape(-1)
        if raw.dtype.kind not in "iu":
            raw = np.asarray(proxies, dtype=object).reshape(-1)
            if any(not (0 <= int(v) < (1 << 32)) for v in raw):
                raise ValueError("Invalid dynamic AABB tree proxy")
        elif raw.size and (int(raw.min()) < 0 or int(raw.max()) >= (1 << 32)):
            raise ValueError("Invalid dynamic AABB tree proxy")
        return np.ascontiguousarray(raw.astype(np.uint32))

    def destroy_proxies(self, proxies):
        ids = self._ids(proxies)
        _check(library().physgpu_tree_destroy_proxies(self._pipe._ctx, len(ids), ids.ctypes.data))

    def destroy_proxy(self, proxy):
        self.destroy_proxies([proxy])

    def update_proxies(self, proxies, bounds):
        """updateProxy per row; bool array, True where reinserted."""
        ids = self._ids(proxies)
        b, bp = self._aabbs(bounds)
        if len(ids) != len(b):
            raise ValueError("one bounds row per proxy")
        flags = np.zeros(max(len(ids), 1), np.uint8)
        _check(library().physgpu_tree_update_proxies(self._pipe._ctx, len(ids), ids.ctypes.data,
                                                      bp, flags.ctypes.data))
        return flags[:len(ids)].astype(bool)

    def update_proxy(self, proxy, bounds):
        """True when the proxy is reinserted (its bounds left the fat bounds,
        or the fat bounds are more than 4x padding too large)."""
        return bool(self.update_proxies([proxy], [bounds])[0])

    def find_candidate_pairs(self, stats=None):
        """Sorted unique (smaller, larger) user-index pairs, uint64 [P, 2]."""
        L = library()
        n = C.c_uint64(0)
        st = abi.BroadPhaseStats()
        _check(L.physgpu_tree_find_pairs(self._pipe._ctx, self.policy, None, 0, C.byref(n),
                                         C.byref(st)))
        out = np.zeros((max(n.value, 1), 2), np.uint64)
        _check(L.physgpu_tree_find_pairs(self._pipe._ctx, self.policy, out.ctypes.data,
                                         n.value, C.byref(n), C.byref(st)))
        if stats is not None:
            stats.clear()
            stats.update({f: getattr(st, f) for f, _ in abi.BroadPhaseStats._fields_})
        return out[:n.value]

    def size(self):
        live = C.c_uint32(0)
        _check(library().physgpu_tree_size(self._pipe._ctx, C.byref(live)))
        return live.value
