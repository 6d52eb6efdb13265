"""The CPU suite again, against the AddressSanitizer + UBSan builds of the
host library and the oracle (make -C gapbs_amd sanitize, make -C oracle
sanitize; GAPBS_CPU_LIB_SUFFIX=.asan selects them).  The sanitizer runtime
is preloaded into a child interpreter, with libstdc++ so that its
exception interceptors resolve.  Any report aborts the child."""

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBS = [os.path.join(ROOT, "gapbs_amd", "libgaphost.asan.so"), os.path.join(ROOT, "oracle", "liboracle.asan.so")]


def _runtime(name):
    try:
        out = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True, check=True)
    except (OSError, subprocess.CalledProcessError):
        return None
    path = out.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.mark.skipif(not all(os.path.exists(p) for p in LIBS), reason="sanitizer builds not made")
def test_cpu_suite_under_address_sanitizer():
    asan, stdcxx = _runtime("libasan.so"), _runtime("libstdc++.so")
    if not asan or not stdcxx:
        pytest.skip("g++ sanitizer runtime not found")
    env = dict(os.environ, LD_PRELOAD=f"{asan} {stdcxx}", GAPBS_CPU_LIB_SUFFIX=".asan",
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    cmd = [sys.executable, "-m", "pytest", os.path.join(ROOT, "tests"), "-q", "-x", "-m", "not gpu",
           "-p", "no:cacheprovider", "--ignore", os.path.abspath(__file__)]
    r = subprocess.run(cmd, env=env, cwd=ROOT, capture_output=True, text=True, timeout=900)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, tail
