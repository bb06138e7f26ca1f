"""The oracle's C++ (host code) under AddressSanitizer and UndefinedBehaviorSanitizer.

The CPU tests that call the oracle library run again in a child process, against
a build of the same sources with -fsanitize=address,undefined (oracle/Makefile,
target `sanitize`); a sanitizer report ends that process with an error, so the
child's exit status is the verdict. Leak checking is off: the Python interpreter
itself keeps allocations alive at exit.
"""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CHECKED = ["tests/test_oracle.py", "tests/test_fuzzy_oracle.py", "tests/test_backend.py"]


def _gcc_runtime(name: str) -> str | None:
    r = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True)
    p = r.stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_oracle_tests_pass_under_asan_and_ubsan():
    asan = _gcc_runtime("libasan.so")
    if asan is None:
        pytest.skip("gcc's AddressSanitizer runtime is not installed")
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "sanitize"], check=True)
    env = dict(os.environ,
               LD_PRELOAD=asan,
               PSQ_ORACLE_LIB=str(ROOT / "oracle" / "_san" / "libpsq_oracle.so"),
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-m", "not gpu",
                        "-p", "no:cacheprovider", *CHECKED],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=1200)
    tail = (r.stdout[-3000:] + "\n" + r.stderr[-3000:]).strip()
    assert r.returncode == 0, tail
    assert "passed" in r.stdout and "AddressSanitizer" not in r.stderr, tail
