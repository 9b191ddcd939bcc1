"""The C restatement (oracle/o2d_oracle.c) under AddressSanitizer and UndefinedBehaviorSanitizer.

oracle/Makefile's "sanitize" target links o2d_oracle.c with oracle/sanitize_driver.c, which calls
every routine on small inputs (edge cases included) with buffers of exactly their stated length.
A sanitizer report makes the program exit non-zero. CPU only; skipped where gcc or the
sanitizer runtimes are missing.
"""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None,
                    reason="gcc/make not available")
def test_oracle_under_asan_and_ubsan():
    b = subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], capture_output=True, text=True)
    if b.returncode != 0 and ("asan" in b.stderr.lower() or "ubsan" in b.stderr.lower()):
        pytest.skip("sanitizer runtime not installed: " + b.stderr[-200:])
    assert b.returncode == 0, b.stderr
    # verify_asan_link_order=0: the process may already preload a library ahead of ASan's
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(ORACLE, "build", "o2d_oracle_sanitize")], capture_output=True,
                       text=True, env=env, timeout=300)
    assert r.returncode == 0 and "sanitize ok" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]
