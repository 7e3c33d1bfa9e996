"""AddressSanitizer + UBSan build of the CPU oracle (host only).

oracle/sanity_main.c calls every oracle entry point on a small confined
two-chain system, then the cell list and forces again on a state with NaN
and 1e30 coordinates. Any out-of-bounds access or undefined behaviour
(signed overflow, out-of-range float to int conversion) aborts the run.
"""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


def test_oracle_under_sanitizers(tmp_path):
    if shutil.which("g++") is None or shutil.which("make") is None:
        pytest.skip("no host compiler")
    build = str(tmp_path / "build")
    out = subprocess.run(["make", "-s", "-C", ORACLE, f"BUILD={build}", f"{build}/oracle_sanity"],
                         capture_output=True, text=True)
    if out.returncode != 0 and "asan" in (out.stderr + out.stdout).lower():
        pytest.skip("sanitizer runtime not available: " + out.stderr[-300:])
    assert out.returncode == 0, out.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="2")
    run = subprocess.run([f"{build}/oracle_sanity"], capture_output=True, text=True, env=env, timeout=300)
    assert run.returncode == 0, (run.stdout + run.stderr)[-3000:]
    assert "runtime error" not in run.stderr, run.stderr[-3000:]
    assert "oracle sanity: rc=0 finite=1" in run.stdout
