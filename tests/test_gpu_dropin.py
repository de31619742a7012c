"""The C++ drop-ins (include/flowcraft/LbmSolver.h, VoxelizeDevice.h) driven with the
reference test's own call sequence plus the mesh -> grid -> solver -> aerodynamic-summary
chain (tests/cpp/dropin_test.cpp), on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "dropin_test")


def test_cpp_dropin_matches_oracle():
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/dropin_test not built: run __graft_entry__.build()")
    print("[gpu-case] cpp drop-in", flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert r.stdout.count("PASS") == 9 and "FAIL" not in r.stdout
