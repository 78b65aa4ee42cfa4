"""The C++17 host interface (include/mpm/mpm_explicit.hpp) compiles against
the headers and links against the engine library and the oracle (CPU); the
test program's parity checks run on the GPU."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "_build", "explicit_solid_test")


def test_cpp_interface_builds():
    r = subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert os.access(BIN, os.X_OK)


@pytest.mark.gpu
def test_cpp_interface_parity():
    assert os.access(BIN, os.X_OK), "build tests/cpp first (__graft_entry__.build())"
    print("running", BIN, flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK")
