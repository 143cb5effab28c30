"""ctypes mirror of include/physgpu.h (record layouts and enums).

The same layouts are used by the HIP library, the CPU oracle and the
reference shim, so numpy views of these records compare field by field.
"""
import ctypes as C

import numpy as np

OK = 0
INVALID_ARGUMENT = 1
CAPACITY_EXCEEDED = 2
DEVICE_ERROR = 3
UNSUPPORTED = 4
EPA_UNRESOLVED = 5

SWEEP_AND_PRUNE = 0
UNIFORM_GRID = 1
DYNAMIC_TREE = 2

SPHERE = 0
BOX = 1
HULL = 2

HOST = 0
DEVICE = 1

PAIR_SPHERE_SPHERE = 0
PAIR_SPHERE_BOX = 1
PAIR_BOX_BOX = 2
PAIR_SPHERE_HULL = 3
PAIR_BOX_HULL = 4
PAIR_HULL_HULL = 5

# EPA polytope of the GPU narrow phase (include/physgpu.h): exact in LDS up
# to these bounds, then continued over at most EPA_MAX_DISTINCT_FACES
# distinct faces with copy counts.
EPA_MAX_FACES = 256
EPA_MAX_HORIZON = 128
EPA_MAX_DISTINCT_FACES = 1024
EPA_MAX_ITERATIONS = 64


class Aabb(C.Structure):
    _fields_ = [("min", C.c_float * 3), ("max", C.c_float * 3)]


class Pair(C.Structure):
    _fields_ = [("first", C.c_uint32), ("second", C.c_uint32)]


class ProxyPair(C.Structure):
    _fields_ = [("first", C.c_uint64), ("second", C.c_uint64)]


TREE_REBUILD = 0
TREE_REFIT = 1


class BroadPhaseStats(C.Structure):
    _fields_ = [
        ("record_build_ms", C.c_double),
        ("record_sort_ms", C.c_double),
        ("sweep_ms", C.c_double),
        ("pair_sort_ms", C.c_double),
        ("x_window_comparisons", C.c_uint64),
        ("aabb_pairs", C.c_uint64),
        ("grid_comparisons", C.c_uint64),
        ("grid_entries", C.c_uint64),
        ("grid_overflow_aabbs", C.c_uint64),
        ("grid_cell_size", C.c_double),
        ("tree_insertions", C.c_uint64),
        ("tree_removals", C.c_uint64),
        ("tree_reinsertions", C.c_uint64),
        ("tree_node_pair_visits", C.c_uint64),
        ("tree_leaf_checks", C.c_uint64),
        ("tree_height", C.c_uint64),
        ("tree_proxy_count", C.c_uint64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class NarrowPhaseStats(C.Structure):
    _fields_ = [
        ("sphere_sphere_candidates", C.c_uint64),
        ("sphere_box_candidates", C.c_uint64),
        ("box_box_candidates", C.c_uint64),
        ("sphere_sphere_contacts", C.c_uint64),
        ("sphere_box_contacts", C.c_uint64),
        ("box_box_contacts", C.c_uint64),
        ("hull_candidates", C.c_uint64),
        ("hull_contacts", C.c_uint64),
        ("gjk_epa_failures", C.c_uint64),
        ("epa_overflows", C.c_uint64),
        ("epa_unresolved", C.c_uint64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class StepStats(C.Structure):
    """Collision fields of phys::StepStats (physgpu_step_stats)."""
    _fields_ = [
        ("broad_phase_ms", C.c_double),
        ("narrow_phase_ms", C.c_double),
        ("broad_phase_collect_ms", C.c_double),
        ("broad_phase_filter_ms", C.c_double),
        ("possible_pairs", C.c_uint64),
        ("candidate_pairs", C.c_uint64),
        ("contact_count", C.c_uint64),
        ("broad_phase", BroadPhaseStats),
        ("narrow_phase", NarrowPhaseStats),
    ]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_[:7]}
        d["broad_phase"] = self.broad_phase.as_dict()
        d["narrow_phase"] = self.narrow_phase.as_dict()
        return d


class Manifold(C.Structure):
    _fields_ = [
        ("first", C.c_uint32),
        ("second", C.c_uint32),
        ("normal", C.c_float * 3),
        ("point_count", C.c_uint32),
        ("anchor_a", (C.c_float * 3) * 4),
        ("anchor_b", (C.c_float * 3) * 4),
        ("penetration", C.c_float * 4),
        ("friction", C.c_float),
        ("restitution", C.c_float),
        ("pair_index", C.c_uint32),
        ("shape_pair", C.c_uint32),
        ("reserved", C.c_uint32 * 2),
    ]


assert C.sizeof(Manifold) == 160

# numpy dtype with the same layout as physgpu_manifold.
MANIFOLD_DTYPE = np.dtype([
    ("first", "<u4"), ("second", "<u4"), ("normal", "<f4", (3,)), ("point_count", "<u4"),
    ("anchor_a", "<f4", (4, 3)), ("anchor_b", "<f4", (4, 3)), ("penetration", "<f4", (4,)),
    ("friction", "<f4"), ("restitution", "<f4"), ("pair_index", "<u4"), ("shape_pair", "<u4"),
    ("reserved", "<u4", (2,)),
])
assert MANIFOLD_DTYPE.itemsize == 160

PAIR_DTYPE = np.dtype([("first", "<u4"), ("second", "<u4")])


class Colliders(C.Structure):
    _fields_ = [
        ("count", C.c_uint32),
        ("pose", C.POINTER(C.c_float)),
        ("shape_type", C.POINTER(C.c_uint8)),
        ("shape_data", C.POINTER(C.c_float)),
        ("hull_range", C.POINTER(C.c_uint32)),
        ("hull_point_count", C.c_uint32),
        ("hull_points", C.POINTER(C.c_float)),
    ]


class NarrowPhaseOptions(C.Structure):
    _fields_ = [
        ("gjk_max_iterations", C.c_uint32),
        ("epa_max_iterations", C.c_uint32),
        ("epa_tolerance", C.c_float),
        ("reserved", C.c_uint32),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("class_block", C.c_uint32 * 3),
        ("class_groups", C.c_uint32 * 3),
        ("epa_groups", C.c_uint32),
        ("query_block", C.c_uint32),
        ("query_groups", C.c_uint32),
        ("sort_items", C.c_uint32),
        ("sphere_hull_per_wave", C.c_uint32),
    ]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if not isinstance(v, int) else v
        return out


class WorldExtras(C.Structure):
    _fields_ = [
        ("body", C.POINTER(C.c_uint32)),
        ("local_pose", C.POINTER(C.c_float)),
        ("friction", C.POINTER(C.c_float)),
        ("restitution", C.POINTER(C.c_float)),
    ]


def fptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def u32ptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_uint32))


def u8ptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_uint8))


def colliders_struct(scene):
    """Build a Colliders struct over a Scene's numpy arrays (kept alive by scene)."""
    c = Colliders()
    c.count = scene.count
    c.pose = fptr(scene.pose)
    c.shape_type = u8ptr(scene.shape_type)
    c.shape_data = fptr(scene.shape_data)
    c.hull_range = u32ptr(scene.hull_range)
    c.hull_point_count = 0 if scene.hull_points is None else scene.hull_points.shape[1]
    c.hull_points = fptr(scene.hull_points)
    return c
