"""physgpu_find_candidate_pairs_noalloc (the allocation-free broad phase,
through the C ABI) against physgpu_find_candidate_pairs on the same bounds:
the same canonical pair list, bit for bit, for every algorithm, on two calls
in a row (the second with fewer bodies, so the fixed working memory holds the
first call's contents when it starts)."""
import ctypes as C

import numpy as np
import pytest

from physgpu import abi, scenes

pytestmark = pytest.mark.gpu

ALGOS = [abi.SWEEP_AND_PRUNE, abi.UNIFORM_GRID, abi.DYNAMIC_TREE]
MAX_BODIES = 65536
MAX_PAIRS = 131072


@pytest.fixture(scope="module")
def torch_dev(physgpu):
    torch = pytest.importorskip("torch")
    return torch


def _noalloc(torch, lib, bounds, algorithm, capacity):
    """One allocation-free call on torch's current stream; (pairs, count)."""
    n = len(bounds)
    dev_bounds = torch.from_numpy(np.ascontiguousarray(bounds, np.float32)).cuda()
    out32 = torch.zeros((max(capacity, 1), 2), dtype=torch.int32, device="cuda")
    count = torch.zeros(1, dtype=torch.int64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    st = lib.physgpu_find_candidate_pairs_noalloc(
        C.c_void_p(dev_bounds.data_ptr()), C.c_uint32(n), C.c_int(algorithm), C.c_double(0.0),
        C.c_void_p(out32.data_ptr()), C.c_uint64(capacity), C.c_void_p(count.data_ptr()),
        C.c_void_p(stream))
    assert st == abi.OK, lib.physgpu_last_error()
    torch.cuda.synchronize()
    total = int(count.cpu().numpy()[0])
    pairs = out32.cpu().numpy().view(np.uint32)[:min(total, capacity)]
    return pairs, total


@pytest.mark.parametrize("algo", ALGOS)
def test_noalloc_matches_find_candidate_pairs(physgpu, torch_dev, algo):
    lib = physgpu.library()
    lib.physgpu_find_candidate_pairs_noalloc.argtypes = [
        C.c_void_p, C.c_uint32, C.c_int, C.c_double, C.c_void_p, C.c_uint64, C.c_void_p,
        C.c_void_p]
    lib.physgpu_find_candidate_pairs_noalloc.restype = C.c_int
    pipe = physgpu.CollisionPipeline(MAX_BODIES, 1 << 20)
    try:
        for n, seed in ((MAX_BODIES - 1000, 41), (20000, 42)):
            print(f"noalloc algo {algo} n {n}", flush=True)
            sc = scenes.mixed(n, seed=seed, side=scenes.mixed_side_for(n, 1.5), with_hulls=False)
            pipe.set_colliders(sc)
            bounds = pipe.compute_aabbs()
            want = pipe.find_candidate_pairs(bounds, algorithm=algo)
            assert 1000 < len(want) < MAX_PAIRS
            got, total = _noalloc(torch_dev, lib, bounds, algo, MAX_PAIRS)
            assert total == len(want)
            assert np.array_equal(got, want)
            # A small output capacity: the first pairs only, the full count.
            got_small, total_small = _noalloc(torch_dev, lib, bounds, algo, 100)
            assert total_small == len(want)
            assert np.array_equal(got_small, want[:100])
    finally:
        pipe.close()


def test_noalloc_rejects_too_many_bodies(physgpu):
    lib = physgpu.library()
    lib.physgpu_find_candidate_pairs_noalloc.restype = C.c_int
    count = C.c_uint64(0)
    st = lib.physgpu_find_candidate_pairs_noalloc(
        C.c_void_p(1), C.c_uint32(MAX_BODIES + 1), C.c_int(abi.UNIFORM_GRID), C.c_double(0.0),
        None, C.c_uint64(0), C.byref(count), None)
    assert st == abi.INVALID_ARGUMENT
    assert b"PHYSGPU_NOALLOC_MAX_BODIES" in lib.physgpu_last_error()
