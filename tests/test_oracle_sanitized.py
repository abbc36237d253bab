"""The oracle's C code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU): the
sanitized build of the same sources (`make -C oracle sanitize`, gcc) runs every path once
(tests/oracle_san_cases.py) in a child process with libasan preloaded; any invalid access,
use after free or undefined behaviour ends it with a report."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_under_asan_and_ubsan():
    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not asan or not os.path.isabs(asan) or not os.path.exists(asan):
        pytest.skip("no libasan")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    lib = os.path.join(ROOT, "oracle", "_san", "liboracle.so")
    env = dict(os.environ, LD_PRELOAD=asan, ARP_ORACLE_LIB=lib, OMP_NUM_THREADS="2",
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "oracle_san_cases.py")], env=env,
                       capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "all cases ok" in r.stdout
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
