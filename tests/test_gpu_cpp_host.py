"""Runs the C++ host API test (tests/cpp/test_host_api.cpp) on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(__file__), "cpp", "_build", "test_host_api")


def test_cpp_host_api(cuda):
    if not os.path.exists(BIN):
        pytest.fail("C++ host test not built (run __graft_entry__.build())")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=100)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
