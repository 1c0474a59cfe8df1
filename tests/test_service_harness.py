"""The build services' host code (csrc/service.cpp with csrc/pinned_pool.cpp and
csrc/build_result.cpp) under ThreadSanitizer and AddressSanitizer, on the CPU, against stand-in
builders (tests/service_harness.cpp): results reach the right sessions only after their copies
have finished, pinned blocks are not reused or freed while a copy writes them, and failures,
destroyed sessions, slot changes and a tight memory budget leave the rest intact."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "octree-cuda_amd", "csrc")
SOURCES = [os.path.join(CSRC, f) for f in ("service.cpp", "pinned_pool.cpp", "build_result.cpp", "host_copy.cpp", "log.cpp")] + \
          [os.path.join(REPO, "tests", "service_harness.cpp")]

# clang's ThreadSanitizer: g++ 11's misses the unlock inside a timed condition-variable wait
# (pthread_cond_clockwait) and reports a double lock there
ROCM_CLANG = "/opt/rocm/llvm/bin/clang++"
CXX = ROCM_CLANG if os.path.exists(ROCM_CLANG) else shutil.which("clang++")

pytestmark = pytest.mark.skipif(CXX is None, reason="no clang++ with sanitizers")


@pytest.mark.parametrize("sanitizer", ["thread", "address"])
def test_services_under_sanitizer(tmp_path, sanitizer):
    exe = str(tmp_path / f"harness_{sanitizer}")
    cmd = [CXX, "-std=c++17", "-O1", "-g", f"-fsanitize={sanitizer}", "-fno-omit-frame-pointer", "-pthread",
           "-I", CSRC, *SOURCES, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 second_deadlock_stack=1",
               ASAN_OPTIONS="detect_leaks=1 halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-6000:]
    assert "harness: 0 failures" in r.stderr
    assert "WARNING: ThreadSanitizer" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
