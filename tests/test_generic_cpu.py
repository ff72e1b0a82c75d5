"""Host-side check of the runtime-size operators' element stages
(nektar_amd/csrc/generic_kernels.hpp) before they run on a GPU: the same
code compiled as C++ with AddressSanitizer and UndefinedBehaviorSanitizer,
one work-item running every stage in order, on anisotropic hex and tet
cases (regular and deformed geometry, every operator), against the oracle's
per-direction operators to 1e-12.  An index outside any intermediate
buffer, table or index list aborts the harness."""
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
def test_runtime_size_stages_in_bounds_and_equal_to_oracle():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "harness")
        cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
               "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + CSRC,
               os.path.join(ROOT, "tests", "cpu_generic", "harness.cpp"), os.path.join(CSRC, "generic.cpp"),
               os.path.join(CSRC, "basis.cpp"), "-L" + ORACLE, "-loracle", "-Wl,-rpath," + ORACLE, "-o", exe]
        subprocess.run(cmd, check=True, capture_output=True, timeout=300)
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
        r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
        assert r.returncode == 0, r.stdout + r.stderr
        assert r.stdout.startswith("OK"), r.stdout
