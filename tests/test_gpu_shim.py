"""The C++ face (include/GooseFEM/GooseFEM.h) used like the reference's own C++ API: run the
compiled drop-in test (tests/cpp/test_shim.cpp) on the GPU."""

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_shim")


def test_cpp_dropin():
    assert os.path.exists(BIN), "build it first: __graft_entry__.build() (make -C tests/cpp)"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "ALL OK" in r.stdout
