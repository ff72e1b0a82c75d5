"""Host-side check of the matrix-core (MFMA) hex Helmholtz body
(nektar_amd/csrc/hex_mfma_block.hpp) before it runs on a GPU: the same code
compiled as C++ with AddressSanitizer and UndefinedBehaviorSanitizer, 256
threads playing one workgroup's work-items and the 16x16x4 FP64 and FP32
MFMA products emulated per 64-thread wave with the lane layouts measured on
gfx950.  Every order NUMMODES 5..10 with Q = NUMMODES + 1 or + 2, 1, 2 and
4 elements per workgroup where they fit, regular and deformed geometry,
against the oracle's Helmholtz operator (FP64 1e-12, FP32 2e-5 relative).
An index outside the LDS slabs, tables, geometry or element arrays aborts
the harness."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "nektar_amd", "csrc")
ORACLE = os.path.join(ROOT, "oracle")


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isfile(os.path.join(ORACLE, "liboracle.so")),
                    reason="needs g++ and the built oracle")
def test_mfma_helmholtz_in_bounds_and_equal_to_oracle():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "harness")
        cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
               "-fno-omit-frame-pointer", "-pthread", "-Wno-unknown-pragmas", "-I" + CSRC,
               os.path.join(ROOT, "tests", "cpu_mfma", "harness.cpp"), os.path.join(CSRC, "basis.cpp"),
               "-L" + ORACLE, "-loracle", "-Wl,-rpath," + ORACLE, "-o", exe]
        subprocess.run(cmd, check=True, capture_output=True, timeout=600)
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
        r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
        assert r.returncode == 0, r.stdout + r.stderr
        assert r.stdout.strip().splitlines()[-1].startswith("OK"), r.stdout
