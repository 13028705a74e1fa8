"""CPU check of the voxelizer's windowed t evaluation (pv2013_amd/csrc/
voxel_window.h, the code the kernel runs): tests/window_check.hip is built
for the host with hipcc and compares, on a million random and adversarial
(column, facet) pairs, the grid points it flips with the literal expression
evaluated at every point of the column (ComplexGeometry.cpp:315-320)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_window_matches_literal_evaluation(tmp_path):
    exe = str(tmp_path / "window_check")
    subprocess.check_call([HIPCC, "-std=c++17", "-O2", "-ffp-contract=off", "--offload-arch=gfx950",
                           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                           "-I", os.path.join(ROOT, "pv2013_amd", "csrc"),
                           os.path.join(ROOT, "tests", "window_check.hip"), "-o", exe],
                          stderr=subprocess.DEVNULL)
    out = subprocess.run([exe, "1000000"], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    m = re.search(r"pairs (\d+) mismatches (\d+) nontrivial (\d+)", out.stdout)
    assert m and int(m.group(2)) == 0 and int(m.group(3)) > 200000
