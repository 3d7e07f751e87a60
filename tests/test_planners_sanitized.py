"""The host-side stamp planners (plan_stamps_rc, plan_units,
plan_stamp_needs) built with AddressSanitizer and UndefinedBehaviorSanitizer
and driven on the CPU with random stamp lists, including strokes that
straddle canvas column/row 0, checked against their invariants
(tests/cpp/plan_fuzz.cpp, tests/cpp/sanitize_planners.sh).  No device is
used."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCRIPT = os.path.join(ROOT, "tests", "cpp", "sanitize_planners.sh")


@pytest.mark.skipif(not os.path.exists("/opt/rocm/bin/hipcc") or not shutil.which("bash"),
                    reason="needs hipcc to build the sanitized library")
def test_planners_under_asan_ubsan():
    r = subprocess.run(["bash", SCRIPT, "1500"], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "0 failures" in r.stdout, r.stdout
