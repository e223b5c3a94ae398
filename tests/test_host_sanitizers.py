"""The library's host code (parameter capture, grid staging and partition,
Becke weights on the host, packed-index maps, ABI argument checks, timers)
built with AddressSanitizer and UBSan, driven by the CPU tests that reach it.

`make -C lio_amd/csrc asan` builds api.cpp, grid.cpp and timers.cpp with
-fsanitize=address,undefined into _build/asan/libg2g.so (the HIP host code of
engine.hip is linked unsanitized: it needs a GPU to run).  The CPU tests of
the host partition, the Fortran ABI link contract, the timers and the
server's host logic then run against that build in a child process; any
sanitizer report fails the test.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "lio_amd", "csrc")
ASAN_LIB = os.path.join(CSRC, "_build", "asan", "libg2g.so")
RUNTIMES = ["/usr/lib/x86_64-linux-gnu/libasan.so.6", "/usr/lib/x86_64-linux-gnu/libstdc++.so.6"]


def test_host_code_under_address_and_ub_sanitizers():
    if not all(os.path.exists(p) for p in RUNTIMES):
        pytest.skip("gcc's AddressSanitizer runtime is not installed")
    r = subprocess.run(["make", "-C", CSRC, "-j8", "asan"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    env = dict(os.environ, LIO_AMD_LIBRARY=ASAN_LIB, LD_PRELOAD=" ".join(RUNTIMES),
               ASAN_OPTIONS="detect_leaks=0:protect_shadow_gap=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    tests = [os.path.join(ROOT, "tests", f) for f in
             ("test_library.py", "test_abi_timers.py", "test_serve_host.py")]
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-m", "not gpu",
                        "-p", "no:cacheprovider"] + tests,
                       capture_output=True, text=True, env=env, cwd=ROOT, timeout=900)
    out = r.stdout + r.stderr
    print(out[-3000:])
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out
