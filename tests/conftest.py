import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "phys-engine_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libphysgpu.so")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    if not os.path.exists(O.ORACLE_SO):
        O.build(ref=False)
    return O.CpuCollision("oracle")


@pytest.fixture(scope="session")
def reference():
    import oracle as O
    if not O.ref_available():
        pytest.skip("oracle/_ref (the reference build) is not present")
    return O.CpuCollision("ref")


@pytest.fixture(scope="session")
def physgpu():
    # torch ships its own HIP runtime; when torch's CUDA is initialised first,
    # libphysgpu.so binds to that runtime and both work in one process (the
    # other order leaves torch without a device).
    try:
        import torch
        if torch.cuda.is_available():
            torch.zeros(1, device="cuda")
    except ImportError:
        pass
    import physgpu as P
    P.library()  # raises when the HIP library is missing: never a silent fallback
    if P.device_count() < 1:
        raise RuntimeError("gpu test selected but no HIP device is visible")
    return P


def golden(name):
    path = os.path.join(GOLDEN, name)
    return np.load(path, allow_pickle=False)


def words(m):
    """uint32 view of manifold records (40 words each)."""
    return np.ascontiguousarray(m).view(np.uint32).reshape(len(m), 40)


# Words compared bit for bit: normal, count, anchors, penetration, friction,
# restitution (pair bookkeeping words are checked separately).
GEOM_WORDS = slice(2, 36)


def assert_manifolds_equal(got, want, what=""):
    assert len(got) == len(want), f"{what}: {len(got)} manifolds vs {len(want)}"
    if not len(got):
        return
    g, w = words(got), words(want)
    assert np.array_equal(g[:, 0:2], w[:, 0:2]), f"{what}: body pairs differ"
    bad = np.nonzero((g[:, GEOM_WORDS] != w[:, GEOM_WORDS]).any(axis=1))[0]
    if len(bad):
        i = bad[0]
        raise AssertionError(f"{what}: {len(bad)} manifolds differ bitwise; first at {i}:\n"
                             f"got  {got[i]}\nwant {want[i]}")


def compact(records, hit):
    """Per-pair oracle slots -> compacted manifolds in candidate order."""
    return records[np.asarray(hit, bool)]
