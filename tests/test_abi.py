"""The C-ABI library loads on a CPU-only host and exports every entry point that
include/physgpu.h declares (no compute calls without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "physgpu.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(physgpu_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("physgpu_create", "physgpu_find_candidate_pairs", "physgpu_generate_contacts",
                 "physgpu_collide", "physgpu_set_world", "physgpu_compute_aabbs"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import physgpu
    lib = physgpu.library()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_device_count_without_gpu_is_safe():
    import physgpu
    assert physgpu.device_count() >= 0


def test_create_rejects_bad_arguments_without_touching_a_device():
    import physgpu
    lib = physgpu.library()
    ctx = ctypes.c_void_p()
    assert lib.physgpu_create(ctypes.byref(ctx), 0, 0, 16, 0) == 1  # zero colliders
    assert b"max_colliders" in lib.physgpu_last_error()
    assert lib.physgpu_create(None, 0, 16, 16, 0) == 1


def test_reference_interface_names():
    import physgpu
    assert int(physgpu.BroadPhaseAlgorithm.SweepAndPrune) == 0
    assert int(physgpu.BroadPhaseAlgorithm.UniformGrid) == 1
    assert int(physgpu.BroadPhaseAlgorithm.DynamicTree) == 2
    try:
        physgpu.BroadPhase.find_candidate_pairs([[0, 0, 0, 1, 1, 1]], None,
                                                physgpu.BroadPhaseAlgorithm.DynamicTree)
    except ValueError as e:
        assert "DynamicTree" in str(e)
    else:
        raise AssertionError("stateless DynamicTree must raise like the reference")
