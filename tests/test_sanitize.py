"""The library's host code under AddressSanitizer + UndefinedBehaviorSanitizer and under
ThreadSanitizer, on the CPU: rsg_host.cpp is compiled with g++ against host-memory stand-ins
for the HIP runtime (tests/sanitize/hip_host_stub.cpp) and driven through set-up (its
multi-threaded reordering, class keys and LDS plans), source and sensor tables, stepping
bookkeeping, state round trips, energy, error paths and batches
(tests/sanitize/sanitize_driver.cpp).  Any sanitizer report, failed expectation or device
allocation left at exit fails the test."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOURCES = [os.path.join(ROOT, "pannde_amd", "csrc", "rsg_host.cpp"),
           os.path.join(ROOT, "tests", "sanitize", "hip_host_stub.cpp"),
           os.path.join(ROOT, "tests", "sanitize", "sanitize_driver.cpp")]


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.exists("/opt/rocm/include/hip/hip_runtime_api.h"),
                    reason="needs g++ and the HIP headers")
@pytest.mark.timeout(600)
@pytest.mark.parametrize("sanitizer", ["address,undefined", "thread"])
def test_host_code_under_sanitizer(tmp_path, sanitizer):
    exe = str(tmp_path / "driver")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=" + sanitizer,
           "-fno-sanitize-recover=undefined", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"] + SOURCES + \
          ["-o", exe, "-lpthread"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300, env=env)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-6000:]
    assert "Sanitizer" not in r.stdout and "runtime error" not in r.stdout, r.stdout[-6000:]
    assert "sanitize driver: 0 failures" in r.stdout
