"""Runs the C++ client written against include/nektar_amd/MatrixFreeOps.hpp
(reference-style MatrixFree/Collections interface) on the GPU; it checks
itself against the oracle (FP64, <= 1e-12)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_mf_cxx.bin")


def test_cxx_client():
    assert os.path.exists(BIN), "build() compiles tests/cpp/test_mf_cxx.bin"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("OK")
