"""The library's host paths under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/cpp/Makefile.sanitize builds a host-only executable from the PCD reader
and writer (pcd.cpp), the error helpers, the self-check hooks (selfcheck.hip:
the device geometry code run on the host) and the CPU oracle (oracle.c), all
with -fsanitize=address,undefined and no recovery; tests/cpp/sanitize_host.cpp
drives them over the golden PCD files, malformed headers, random and
tie-heavy clouds.  Any sanitizer finding or wrong answer fails the test.
"""
import glob
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(not os.path.exists("/opt/rocm/llvm/bin/clang++"), reason="ROCm clang not installed")
def test_host_paths_under_asan_and_ubsan(tmp_path):
    out = tmp_path / "build"
    mk = subprocess.run(["make", "-s", "-j4", "-f", os.path.join(ROOT, "tests", "cpp", "Makefile.sanitize"),
                         f"OUT={out}"], capture_output=True, text=True)
    assert mk.returncode == 0, mk.stderr[-4000:]
    files = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "*.pcd")))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0", UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([str(out / "sanitize_host"), str(tmp_path)] + files, capture_output=True, text=True,
                         env=env, timeout=600)
    assert run.returncode == 0, (run.stdout[-2000:], run.stderr[-6000:])
    assert "OK" in run.stdout
