"""The host-side code built with AddressSanitizer + UndefinedBehaviorSanitizer
(`make -C oracle sanitize`, the reference's Debug flags,
src/CMakeLists.txt:27,39-42): the drop-in's host primitives and taxonomy
store and the C oracle run tests/cpp/sanitize_check.cpp with no report."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_code_is_sanitizer_clean(tmp_path):
    if not shutil.which("g++") or not shutil.which("gcc"):
        pytest.skip("no host compiler")
    p = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"],
                       capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    p = subprocess.run([os.path.join(ROOT, "oracle", "_san", "sanitize_check"), str(tmp_path)],
                       capture_output=True, text=True, timeout=600, env=env)
    assert p.returncode == 0, (p.stdout + p.stderr)[-3000:]
    assert p.stdout.strip() == "ok"
    assert "runtime error" not in p.stderr and "AddressSanitizer" not in p.stderr
