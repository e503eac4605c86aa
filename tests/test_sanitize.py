"""The host half of liblabinius_gpu under ThreadSanitizer and under AddressSanitizer + UBSan, on
the CPU: tests/native builds the worker pool, the scheduler, the transcript and the prover's and
checker's host orchestration against a stand-in for the device half (tests/native/device_stub.cpp)
and stresses them (tests/native/sanitize_driver.cpp): many callers of the pool at once, twelve
scheduler workers on mixed sizes, the batch prover's phases on its host threads (each batch proof
equal to its solo proof), the checker's decoder on damaged proofs, the transcript's split absorb.
Any sanitizer report or failed check fails the test.
"""
import os
import shutil
import subprocess

import pytest

NATIVE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None,
                                reason="needs g++ and make")


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    out = tmp_path_factory.mktemp("sanitize")
    r = subprocess.run(["make", "-C", NATIVE, "-j2", f"OUT={out}"], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-4000:]
    return out


@pytest.mark.parametrize("kind,scale", [("asan", "4"), ("tsan", "1")])
def test_host_code_under_sanitizer(built, kind, scale):
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([str(built / f"driver_{kind}"), scale], capture_output=True, text=True, timeout=900, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "Sanitizer" not in r.stderr, r.stderr[-4000:]
    assert "batch done, failures 0" in r.stderr
