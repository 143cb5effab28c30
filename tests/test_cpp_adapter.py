"""The C++ host adapter (phys-engine_amd/adapter/physgpu.hpp), which gives C++
callers the reference's own collision interfaces over the C ABI, driven by a
g++-built program in the shape of the reference's collision tests
(tests/cpp/adapter_tests.cpp)."""
import os
import subprocess

import pytest

from conftest import ROOT

BIN = os.path.join(ROOT, "tests", "cpp", "_build", "adapter_tests")


def _binary():
    if not os.path.exists(BIN):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], check=True)
    return BIN


def test_adapter_argument_checks_without_device():
    r = subprocess.run([_binary(), "--host-only"], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failure(s)" in r.stdout


@pytest.mark.gpu
def test_adapter_on_device():
    r = subprocess.run([_binary()], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("ok  ") >= 20
