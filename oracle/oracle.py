"""Python access to the CPU oracle — TEST INFRASTRUCTURE.

Two libraries share one contract (include/physgpu.h record layouts):
  liboracle.so        this repository's CPU restatement (oracle/physcpu.cpp)
  _ref/libphysref.so  the reference engine compiled from its own sources
                      (oracle/Makefile `ref`; present only where it was built)

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module; the product never does.
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "phys-engine_amd"))

from physgpu import abi  # noqa: E402

ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libphysref.so")
REF_ROOT = "/root/reference"


def build(ref=None):
    """Compile the restatement, and the reference shim when the reference
    sources are present (ref=None) or when asked (ref=True)."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    if ref or (ref is None and os.path.isdir(os.path.join(REF_ROOT, "src"))):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def ref_available():
    return os.path.exists(REF_SO)


class CpuCollision:
    """One of the two CPU libraries behind a common numpy interface."""

    def __init__(self, which="oracle"):
        self.which = which
        path = ORACLE_SO if which == "oracle" else REF_SO
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run oracle/Makefile)")
        self.lib = C.CDLL(path)
        p = "orc_" if which == "oracle" else "ref_"
        self.p = p
        L = self.lib
        vp = C.c_void_p
        getattr(L, p + "last_error").restype = C.c_char_p
        getattr(L, p + "compute_aabbs").argtypes = [C.POINTER(abi.Colliders), vp]
        getattr(L, p + "find_candidate_pairs").argtypes = [
            vp, C.c_uint32, C.c_int, C.c_uint32, vp, C.c_uint64, C.POINTER(C.c_uint64),
            C.POINTER(abi.BroadPhaseStats)]
        if which == "oracle":
            L.orc_generate_contacts.argtypes = [
                C.POINTER(abi.Colliders), vp, C.c_uint64, C.c_uint32, C.c_uint32, C.c_float,
                C.c_uint32, vp, vp, vp]
            L.orc_generate_contacts_f64.argtypes = [
                C.POINTER(abi.Colliders), vp, C.c_uint64, C.c_uint32, C.c_uint32, C.c_float,
                vp, vp]
            L.orc_count_flops.argtypes = [
                C.POINTER(abi.Colliders), vp, C.c_uint64, C.c_uint32, C.c_uint32, C.c_float,
                C.c_uint32, vp]
        else:
            L.ref_generate_contacts.argtypes = [C.POINTER(abi.Colliders), vp, C.c_uint64, vp, vp]
        getattr(L, p + "gjk_epa").argtypes = [
            C.POINTER(abi.Colliders), vp, C.c_uint64, C.c_uint32, C.c_uint32, C.c_float,
            vp, vp, vp, vp]
        if which == "oracle":
            L.orc_set_epa_bounds.argtypes = [C.c_uint32, C.c_uint32]
            L.orc_set_epa_bounds.restype = None
            L.orc_set_epa_continuation.argtypes = [C.c_uint32, C.c_int]
            L.orc_set_epa_continuation.restype = None
            L.orc_set_test_mutation.argtypes = [C.c_int]
            L.orc_set_test_mutation.restype = None
            L.orc_contact_adjoint_f64.argtypes = [C.POINTER(abi.Colliders), vp, C.c_uint64, vp,
                                                  vp, vp, C.c_uint32, vp, vp]
            L.orc_contact_fd_f64.argtypes = [C.POINTER(abi.Colliders), vp, C.c_uint64, vp, vp, vp,
                                             C.c_double, C.c_uint32, vp, vp]
            L.orc_contact_geometry_f64.argtypes = [C.POINTER(abi.Colliders), vp, C.c_uint64, vp,
                                                   vp, vp, vp]
            L.orc_contact_choices.argtypes = [C.POINTER(abi.Colliders), vp, C.c_uint64, vp]
        getattr(L, p + "world_collide").argtypes = [
            C.POINTER(abi.Colliders), vp, vp, vp, vp, C.c_int, C.c_uint32, C.c_uint32,
            C.c_uint32, C.c_uint32, vp, C.c_uint64, C.POINTER(C.c_uint64), vp, vp]
        getattr(L, p + "world_collide_bodies").argtypes = [
            C.POINTER(abi.Colliders), vp, vp, vp, vp, vp, C.c_int, C.c_uint32, C.c_uint32,
            C.c_uint32, C.c_uint32, vp, C.c_uint64, C.POINTER(C.c_uint64), vp, vp]

    def set_epa_bounds(self, max_faces=0, max_edges=0):
        """Bounded EPA polytope (the GPU's limits) or 0, 0 for the reference's
        unbounded behaviour."""
        assert self.which == "oracle"
        self.lib.orc_set_epa_bounds(max_faces, max_edges)

    def mirror_gpu_epa(self, on=True, continuation=True):
        """EPA as the GPU narrow phase runs it: the reference's list within
        the LDS bounds, then (with continuation) the multiplicity form over
        at most EPA_MAX_DISTINCT_FACES distinct faces, with the order of new
        faces within an iteration untracked.  on=False restores the
        reference's unbounded EPA."""
        if on:
            self.set_epa_bounds(abi.EPA_MAX_FACES, abi.EPA_MAX_HORIZON)
            self.set_epa_continuation(abi.EPA_MAX_DISTINCT_FACES if continuation else 0,
                                      exact_order=False)
        else:
            self.set_epa_bounds(0, 0)
            self.set_epa_continuation(0)

    def set_epa_continuation(self, multi_faces=0, exact_order=True):
        """Past the polytope bounds, continue EPA over distinct faces with
        multiplicities (at most multi_faces of them; 0: stop with the
        overflow status), as the GPU narrow phase does."""
        assert self.which == "oracle"
        self.lib.orc_set_epa_continuation(multi_faces, int(exact_order))

    def set_test_mutation(self, mutation=0):
        """Plant a known mistake in the oracle's hull contact point (tests of
        the tests only): 1 supports along the wrong signs, 2 A's support
        point instead of the midpoint; 0 restores the oracle."""
        assert self.which == "oracle"
        self.lib.orc_set_test_mutation(mutation)

    def _check(self, rc, what):
        if rc not in (abi.OK,):
            msg = getattr(self.lib, self.p + "last_error")()
            raise RuntimeError(f"{self.which} {what}: status {rc}: {msg!r}")

    def compute_aabbs(self, scene):
        c = abi.colliders_struct(scene)
        out = np.zeros((scene.count, 6), np.float32)
        rc = getattr(self.lib, self.p + "compute_aabbs")(C.byref(c), out.ctypes.data)
        self._check(rc, "compute_aabbs")
        return out

    def find_pairs(self, aabbs, algorithm=abi.SWEEP_AND_PRUNE, workers=1):
        aabbs = np.ascontiguousarray(aabbs, dtype=np.float32).reshape(-1, 6)
        n = aabbs.shape[0]
        count = C.c_uint64(0)
        stats = abi.BroadPhaseStats()
        fn = getattr(self.lib, self.p + "find_candidate_pairs")
        rc = fn(aabbs.ctypes.data, n, algorithm, workers, None, 0, C.byref(count), C.byref(stats))
        if rc not in (abi.OK, abi.CAPACITY_EXCEEDED):
            self._check(rc, "find_candidate_pairs")
        out = np.zeros((max(count.value, 1), 2), np.uint32)
        rc = fn(aabbs.ctypes.data, n, algorithm, workers, out.ctypes.data, count.value,
                C.byref(count), C.byref(stats))
        self._check(rc, "find_candidate_pairs")
        return out[:count.value], stats.as_dict()

    def generate_contacts(self, scene, pairs, gjk_iterations=64, epa_iterations=64,
                          tolerance=1e-6, workers=1):
        """One manifold record per pair (not compacted) and the hit flags."""
        c = abi.colliders_struct(scene)
        pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
        n = pairs.shape[0]
        out = np.zeros(max(n, 1), abi.MANIFOLD_DTYPE)
        hit = np.zeros(max(n, 1), np.uint8)
        if self.which == "oracle":
            status = np.zeros(max(n, 1), np.int32)
            rc = self.lib.orc_generate_contacts(C.byref(c), pairs.ctypes.data, n, gjk_iterations,
                                                epa_iterations, tolerance, workers,
                                                out.ctypes.data, hit.ctypes.data,
                                                status.ctypes.data)
        else:
            rc = self.lib.ref_generate_contacts(C.byref(c), pairs.ctypes.data, n,
                                                out.ctypes.data, hit.ctypes.data)
        self._check(rc, "generate_contacts")
        return out[:n], hit[:n].astype(bool)

    def generate_contacts_f64(self, scene, pairs, gjk_iterations=64, epa_iterations=64,
                              tolerance=1e-6):
        assert self.which == "oracle"
        c = abi.colliders_struct(scene)
        pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
        n = pairs.shape[0]
        out = np.zeros((max(n, 1), 32), np.float64)
        hit = np.zeros(max(n, 1), np.uint8)
        rc = self.lib.orc_generate_contacts_f64(C.byref(c), pairs.ctypes.data, n, gjk_iterations,
                                                epa_iterations, tolerance, out.ctypes.data,
                                                hit.ctypes.data)
        self._check(rc, "generate_contacts_f64")
        return out[:n], hit[:n].astype(bool)

    def count_flops(self, scene, pairs, gjk_iterations=64, epa_iterations=64, tolerance=1e-6,
                    workers=1):
        """Floating-point operations of each pair's contact in FP32, uint64 [n, 8]:
        add/subtract, multiply, divide, square root, of which in GJK, GJK
        iterations, EPA iterations, touching (physcpu.cpp orc_count_flops)."""
        assert self.which == "oracle"
        c = abi.colliders_struct(scene)
        pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
        n = pairs.shape[0]
        out = np.zeros((max(n, 1), 8), np.uint64)
        rc = self.lib.orc_count_flops(C.byref(c), pairs.ctypes.data, n, gjk_iterations,
                                      epa_iterations, tolerance, workers, out.ctypes.data)
        self._check(rc, "count_flops")
        return out[:n]

    def gjk_epa(self, scene, pairs, gjk_iterations=64, epa_iterations=64, tolerance=1e-6):
        c = abi.colliders_struct(scene)
        pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
        n = pairs.shape[0]
        status = np.zeros(max(n, 1), np.int32)
        normal = np.zeros((max(n, 1), 3), np.float32)
        depth = np.zeros(max(n, 1), np.float32)
        simplex = np.zeros((max(n, 1), 13), np.float32)
        rc = getattr(self.lib, self.p + "gjk_epa")(
            C.byref(c), pairs.ctypes.data, n, gjk_iterations, epa_iterations, tolerance,
            status.ctypes.data, normal.ctypes.data, depth.ctypes.data, simplex.ctypes.data)
        self._check(rc, "gjk_epa")
        return status[:n], normal[:n], depth[:n], simplex[:n]

    def world_collide(self, scene, algorithm=abi.SWEEP_AND_PRUNE, bp_workers=1, np_workers=1,
                      warmup=0, steps=1, capacity=None):
        """World-level collision section; returns (manifolds, timing dict, np stats).

        With scene.body set, colliders with equal body values share one body
        (body_position, friction and restitution then hold each collider's
        body's values); manifold first / second are collider indices in the
        oracle and body indices in the reference (ContactManifold::bodyA / B)."""
        c = abi.colliders_struct(scene)
        count = C.c_uint64(0)
        timing = np.zeros(6, np.float64)
        np_stats = np.zeros(6, np.uint64)
        ptr = (lambda a: None if a is None else a.ctypes.data)
        cap = capacity if capacity is not None else 4 * scene.count + 16
        out = np.zeros(cap, abi.MANIFOLD_DTYPE)
        common = (algorithm, bp_workers, np_workers, warmup, steps, out.ctypes.data, cap,
                  C.byref(count), timing.ctypes.data, np_stats.ctypes.data)
        body = getattr(scene, "body", None)
        if body is not None:
            body = np.ascontiguousarray(body, np.uint32)
            fn = getattr(self.lib, self.p + "world_collide_bodies")
            rc = fn(C.byref(c), ptr(scene.body_position), body.ctypes.data, ptr(scene.local_pose),
                    ptr(scene.friction), ptr(scene.restitution), *common)
        else:
            fn = getattr(self.lib, self.p + "world_collide")
            rc = fn(C.byref(c), ptr(scene.body_position), ptr(scene.local_pose),
                    ptr(scene.friction), ptr(scene.restitution), *common)
        self._check(rc, "world_collide")
        t = dict(broadphase_ms=timing[0], narrowphase_ms=timing[1], aabb_ms=timing[2],
                 total_ms=timing[3], candidate_pairs=int(timing[4]), possible_pairs=timing[5])
        return out[:count.value], t, np_stats

    # ---- narrow-phase adjoint (FP64): pairs are the manifolds' (first, second),
    # upstream gradients SoA planes g_normal [3, m], g_depth [4, m], g_point [12, m].
    @staticmethod
    def _planes(a, rows, m):
        if a is None:
            return None
        a = np.ascontiguousarray(a, np.float32)
        assert a.shape == (rows, m), (a.shape, rows, m)
        return a

    def contact_adjoint_f64(self, scene, pairs, g_normal=None, g_depth=None, g_point=None,
                            workers=8):
        """Dual-number (forward-mode) Jacobian-vector products: grad [14, m]
        (A's px..qw then B's) and status [m] (1 done, 0 no contact, -1 hull)."""
        c = abi.colliders_struct(scene)
        pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        m = pairs.shape[0]
        gn, gd, gp = (self._planes(g_normal, 3, m), self._planes(g_depth, 4, m),
                      self._planes(g_point, 12, m))
        grad = np.zeros((14, max(m, 1)), np.float64)
        status = np.zeros(max(m, 1), np.int32)
        ptr = (lambda a: None if a is None else a.ctypes.data)
        rc = self.lib.orc_contact_adjoint_f64(C.byref(c), pairs.ctypes.data, m, ptr(gn), ptr(gd),
                                              ptr(gp), workers, grad.ctypes.data,
                                              status.ctypes.data)
        self._check(rc, "contact_adjoint_f64")
        return grad[:, :m], status[:m]

    def contact_fd_f64(self, scene, pairs, g_normal=None, g_depth=None, g_point=None, h=1e-6,
                       workers=8):
        """FP64 central differences of the same linearised loss: grad [14, m]
        and stable [m] (1 when no discrete choice changed under the steps)."""
        c = abi.colliders_struct(scene)
        pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        m = pairs.shape[0]
        gn, gd, gp = (self._planes(g_normal, 3, m), self._planes(g_depth, 4, m),
                      self._planes(g_point, 12, m))
        grad = np.zeros((14, max(m, 1)), np.float64)
        stable = np.zeros(max(m, 1), np.int32)
        ptr = (lambda a: None if a is None else a.ctypes.data)
        rc = self.lib.orc_contact_fd_f64(C.byref(c), pairs.ctypes.data, m, ptr(gn), ptr(gd),
                                         ptr(gp), h, workers, grad.ctypes.data,
                                         stable.ctypes.data)
        self._check(rc, "contact_fd_f64")
        return grad[:, :m], stable[:m].astype(bool)

    def contact_geometry_f64(self, scene, pairs):
        """FP64 normal [3, m], depth [4, m], world points [12, m], count [m]."""
        c = abi.colliders_struct(scene)
        pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        m = pairs.shape[0]
        normal = np.zeros((3, max(m, 1)), np.float64)
        depth = np.zeros((4, max(m, 1)), np.float64)
        point = np.zeros((12, max(m, 1)), np.float64)
        count = np.zeros(max(m, 1), np.int32)
        rc = self.lib.orc_contact_geometry_f64(C.byref(c), pairs.ctypes.data, m,
                                               normal.ctypes.data, depth.ctypes.data,
                                               point.ctypes.data, count.ctypes.data)
        self._check(rc, "contact_geometry_f64")
        return normal[:, :m], depth[:, :m], point[:, :m], count[:m]

    CHOICE_FIELDS = ("hit", "sel", "flip", "fallback", "count", "branch", "face", "clamped",
                     "support_a", "support_b", "zero_depths", "from_b")

    def contact_choices(self, scene, pairs):
        """Discrete choices of each pair's FP64 contact (dict of int arrays)."""
        c = abi.colliders_struct(scene)
        pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        m = pairs.shape[0]
        out = np.zeros((max(m, 1), 12), np.int32)
        rc = self.lib.orc_contact_choices(C.byref(c), pairs.ctypes.data, m, out.ctypes.data)
        self._check(rc, "contact_choices")
        return {f: out[:m, i] for i, f in enumerate(self.CHOICE_FIELDS)}
