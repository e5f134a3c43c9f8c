"""Host-side checks of device helpers that are also callable on the CPU (no GPU)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "particle-fluid-simulation_amd", "csrc")


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="hipcc is needed to compile the header")
def test_row3_bounds_exhaustive(tmp_path):
    """The row lookup's in-grid test (a wrapped form of it once sent a load 2^32 - 1 entries
    past the halo grid) against the exact integer condition, for every row length up to 512
    and every middle cell within 2000 of the row plus the int32 extremes."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    exe = tmp_path / "row3_check"
    src = os.path.join(ROOT, "tests", "native", "row3_check.cpp")
    subprocess.run([hipcc, "-x", "hip", "--offload-arch=gfx950", "-O1", "-std=c++17", "-I", CSRC,
                    src, "-o", str(exe)], check=True, capture_output=True, timeout=300)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout
    assert "mismatches 0" in out.stdout
