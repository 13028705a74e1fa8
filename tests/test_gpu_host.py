"""lbm_run end to end: STL + configuration -> GPU voxelizer -> steps -> VTU
files, against orc_run (the CPU oracle driven by the same configuration file)
file for file."""
import os
import re
import subprocess

import numpy as np
import pytest

from host_case import write_case
from test_oracle_driver import case_with
from vtu import read_vtu

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "pv2013_amd", "csrc", "lbm_run")
ORC_RUN = os.path.join(ROOT, "oracle", "orc_run")


def run(exe, cfg, timeout=300):
    r = subprocess.run([exe, cfg], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def same_outputs(a, b, iters, points_files=()):
    """Every VTU array of iteration n bit for bit, and the boundary-point files
    byte for byte, between the output folders a and b."""
    for n in iters:
        for f in ("Fluid_Variables/%d_rank_0.vtu" % n, "Wall_Shear_Stress_VTK/%d_rank_0_wall_shear_stress.vtu" % n):
            x, y = read_vtu(os.path.join(a, f)), read_vtu(os.path.join(b, f))
            assert sorted(x) == sorted(y), f
            for k in x:
                assert x[k].dtype == y[k].dtype and np.array_equal(x[k].view(np.uint8), y[k].view(np.uint8)), (f, k)
    for f in points_files:
        assert open(os.path.join(a, f), "rb").read() == open(os.path.join(b, f), "rb").read(), f


@pytest.mark.parametrize("fmt", ["ascii", "raw"])
def test_lbm_run_matches_orc_run(tmp_path, fmt):
    """One case file drives the GPU engine (lbm_run) and the CPU oracle
    (orc_run): FP64 with the reference's quirks, a steady parabolic inlet, a
    pressure outlet; the files of iterations 0, 50 and 100 (Velocity, Density,
    WSS of the state each iteration collides, Pressure after it) and the
    boundary-point files agree bit for bit."""
    d = str(tmp_path)
    cfg, _, _ = write_case(d, steps=101, every=50)
    gpu = case_with(os.path.join(d, "gpu.cfg"), cfg, output_dir=os.path.join(d, "gpu"), output_format=fmt,
                    boundary_points=1)
    cpu = case_with(os.path.join(d, "cpu.cfg"), cfg, output_dir=os.path.join(d, "cpu"), output_format=fmt,
                    boundary_points=1)
    print("lbm_run / orc_run", fmt, flush=True)
    run(RUN, gpu)
    run(ORC_RUN, cpu)
    same_outputs(os.path.join(d, "gpu"), os.path.join(d, "cpu"), (0, 50, 100),
                 ["%sboundaryPointsFromGeometry.vtk" % p for p in ("inlet", "outlet", "Wall")])


def test_lbm_run_autotune_gives_identical_files(tmp_path):
    """`autotune 1` in the configuration times the candidate layouts first and
    runs the fastest; every layout gives the same bits, so the VTU files equal
    those of the untuned run byte for byte.  A second run reads the choice
    back from the tuning cache."""
    cfg, _, _ = write_case(str(tmp_path), storage="f32")
    base = open(cfg).read()
    tuned_cfg = os.path.join(str(tmp_path), "tuned.cfg")
    cache = os.path.join(str(tmp_path), "tune.cache")
    with open(tuned_cfg, "w") as fh:
        fh.write(base.replace(os.path.join(str(tmp_path), "out"), os.path.join(str(tmp_path), "out_tuned")))
        fh.write("autotune 1\nautotune_cache %s\n" % cache)
    print("lbm_run plain and tuned", flush=True)
    r = subprocess.run([RUN, cfg], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    for attempt in range(2):
        r = subprocess.run([RUN, tuned_cfg], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        m = re.search(r"autotune: .*\((\d+) candidates", r.stdout)
        assert m, r.stdout
        assert (int(m.group(1)) > 0) == (attempt == 0), r.stdout  # the second run reads the cache
        for f in ("Fluid_Variables/100_rank_0.vtu", "Wall_Shear_Stress_VTK/100_rank_0_wall_shear_stress.vtu"):
            a = open(os.path.join(str(tmp_path), "out", f), "rb").read()
            b = open(os.path.join(str(tmp_path), "out_tuned", f), "rb").read()
            assert a == b, f
