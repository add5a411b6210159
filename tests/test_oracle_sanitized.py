"""The FP64 oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

`make -C oracle sanitized` builds the oracle source with both sanitizers (every finding fatal) into
oracle/_san/.  The oracle tests then run in a child interpreter that loads that build (MPM_ORACLE_LIB)
with the ASan runtime preloaded (the interpreter itself is not instrumented).  A heap overflow, a
use-after-free, a signed overflow, an out-of-range shift or a misaligned access in any oracle entry point
that the tests reach ends the child with a non-zero status.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_tests_pass_under_asan_and_ubsan():
    if shutil.which("g++") is None or shutil.which("make") is None:
        pytest.skip("g++ and make are needed to build the sanitized oracle")
    libasan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(libasan) or not os.path.exists(libasan):
        pytest.skip("no ASan runtime for g++")
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "sanitized"], check=True, stdout=subprocess.DEVNULL)
    env = dict(os.environ,
               LD_PRELOAD=libasan,
               MPM_ORACLE_LIB=os.path.join(ROOT, "oracle", "_san", "liboracle_mpm.so"),
               ASAN_OPTIONS="detect_leaks=0:exitcode=86",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               OMP_NUM_THREADS="4")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                        os.path.join(ROOT, "tests", "test_oracle.py")],
                       env=env, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, (r.stdout[-4000:], r.stderr[-4000:])
    assert " passed" in r.stdout
