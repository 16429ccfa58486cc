"""The served form's scan result ring (include/indexingpu_amd/scan_ring.hpp), built with g++
and checked on the host: rows of completed requests are not reused before the ring wraps
(also across an empty ring), a wave as large as the ring fits an empty ring, rings that
hold a run's rows exactly never reuse a row, and random reserve/release never overlaps a
live reservation."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is not installed")
def test_scan_ring(tmp_path):
    exe = tmp_path / "scan_ring_check"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", os.path.join(REPO, "include"),
                           os.path.join(HERE, "scan_ring", "scan_ring_check.cpp"), "-o", str(exe)])
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip() == "ok"
