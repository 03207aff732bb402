"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: a CPU
build of oracle/sc_oracle.c with a driver that calls every entry point
(oracle/sanitize_driver.c, `make -C oracle sanitize`) must run clean."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None, reason="needs gcc and make")
def test_oracle_runs_clean_under_asan_and_ubsan(tmp_path):
    b = subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "sanitize", f"OUT={tmp_path}"],
                       capture_output=True, text=True)
    assert b.returncode == 0, b.stdout + b.stderr
    env = dict(os.environ, ASAN_OPTIONS="abort_on_error=0:halt_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    r = subprocess.run([str(tmp_path / "sanitize_driver")], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
    assert "all entry points ok" in r.stdout
