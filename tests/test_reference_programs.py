"""The reference's own test programs (src/tests/test_constraints.cpp,
test_topology.cpp, test_pd_constraints.cpp), compiled unchanged against this
repository's headers and linked against its libraries by oracle/Makefile
into oracle/_ref/. Each prints "Results: <p> passed, <f> failed" and exits
non-zero on a failure. Skipped where they were not built (the reference's
sources are needed to build them, not to run them).

test_pd_constraints is built without CUDA_MS_HAVE_CUDA, because no
cuda_runtime.h exists here: its Tests 1-5 then run the GPU solver through the
C++ API but read positions back from host arrays, so of its checks only Test
6 (the CPU reference backend, CpuStretchReferenceSolver) tests results."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")


def run_program(name):
    exe = os.path.join(REF, name)
    if not os.path.exists(exe):
        pytest.skip(f"{name} not built (reference sources absent when oracle/ was built)")
    env = dict(os.environ, CLOTH_QUIET="1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    m = re.search(r"Results: (\d+) passed, (\d+) failed", r.stdout)
    assert m, r.stdout[-2000:] + r.stderr[-2000:]
    return r, int(m.group(1)), int(m.group(2))


@pytest.mark.parametrize("name,least", [("test_constraints", 6), ("test_topology", 164)])
def test_reference_program_passes(name, least):
    r, passed, failed = run_program(name)
    assert r.returncode == 0 and failed == 0 and passed >= least, r.stdout[-2000:]


@pytest.mark.gpu
def test_reference_pd_constraints_program():
    print("reference program test_pd_constraints", flush=True)
    r, passed, failed = run_program("test_pd_constraints")
    assert r.returncode == 0 and failed == 0, r.stdout[-3000:]
    assert "CPU ref Y-range" in r.stdout
