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


SAN = os.path.join(os.path.dirname(__file__), "cpp", "_build", "test_host_sanitize")


def test_host_sanitizer_driver_with_device(cuda):
    """The host-sanitized driver with its device part: operators and
    composites, concurrent H.v from four host threads, the batched engine,
    the serving loop fed and drained by eight host threads with the Lanczos
    lane, an oversize request, the no-allocation Lanczos.  Host code under
    AddressSanitizer and UndefinedBehaviorSanitizer (leak detection off: the
    HIP runtime's allocations live until exit)."""
    if not os.path.exists(SAN):
        pytest.fail("host-sanitizer driver not built (run __graft_entry__.build())")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([SAN], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "ALL PASSED" in r.stdout
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-3000:]
