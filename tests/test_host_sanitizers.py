"""The library's host code under AddressSanitizer and UndefinedBehavior-
Sanitizer (leak detection on): libspbp is rebuilt with the sanitizers on its
host side only (device code as usual) and tests/host/host_check.cpp drives
every C-ABI entry point along the paths a host reaches without a usable GPU
-- argument validation, NULL contexts, bad sizes, the host table and rotation
routines, creation failing at the device, the error buffers.  Any sanitizer
report or failed check fails the test.  About a minute of compilation."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# host-only sanitizers: each -fsanitize applies to the host compilation
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-fno-omit-frame-pointer", "-g",
       "-O1"]


@pytest.mark.skipif(not shutil.which(HIPCC) and not os.path.exists(HIPCC), reason="hipcc not available")
def test_host_code_clean_under_asan_and_ubsan(tmp_path):
    lib = str(tmp_path / "libspbp_san.so")
    exe = str(tmp_path / "host_check")
    inc = os.path.join(ROOT, "include")
    subprocess.check_call([HIPCC, "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared", "-ffp-contract=off",
                           "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", *SAN,
                           "-I", inc, "-o", lib, os.path.join(ROOT, "spider_amd", "csrc", "spbp.hip"),
                           "-L/opt/rocm/lib", "-lhipfft", "-Wl,-rpath,/opt/rocm/lib"],
                          stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    subprocess.check_call([HIPCC, "-std=c++17", *SAN, "-I", inc, "-o", exe, os.path.join(HERE, "host", "host_check.cpp"),
                           "-L" + str(tmp_path), "-lspbp_san", "-Wl,-rpath," + str(tmp_path),
                           "-Wl,-rpath,/opt/rocm/lib"], stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    p = subprocess.run([exe], env=env, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "host checks passed" in p.stdout
    assert "ERROR: AddressSanitizer" not in p.stderr and "runtime error" not in p.stderr, p.stderr
