"""lbm_run end to end: STL + configuration -> GPU voxelizer -> steps -> VTU
files, identical to the same case run through the C ABI from Python."""
import os
import re
import subprocess

import numpy as np
import pytest

from host_case import write_case

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "pv2013_amd", "csrc", "lbm_run")


def vtu_array(path, name):
    text = open(path).read()
    m = re.search(r'Name="%s"[^>]*>(.*?)</DataArray>' % name, text, re.S)
    return np.array([float(v) for v in m.group(1).split()])


def test_lbm_run_matches_python_path(tmp_path):
    import pv2013_amd as P
    cfg, stl, info = write_case(str(tmp_path))
    print("lbm_run", cfg, flush=True)
    r = subprocess.run([RUN, cfg], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = os.path.join(str(tmp_path), "out")
    for f in ("Fluid_Variables/0_rank_0.vtu", "Fluid_Variables/100_rank_0.vtu",
              "Wall_Shear_Stress_VTK/100_rank_0_wall_shear_stress.vtu"):
        assert os.path.exists(os.path.join(out, f)), f
    ctx = P.Context(0)
    tri = P.read_stl(stl)
    dims = info["dims"]
    g = P.make_grid((0, 0, 0), 1.0, dims, (0, dims[0] - 1, 0, dims[1] - 1, 0, dims[2] - 1))
    flags, _ = ctx.voxelize(tri, g)
    flags, _ = ctx.classify(g, flags, [P.make_plane(2, "1"), P.make_plane(2, str(info["outlet_z"]))])
    cx, cy = info["inlet_center"]
    bcs = [P.make_bc(P.BC_VELOCITY, 2, 1, u=(0, 0, 0.03), profile=P.PROFILE_PARABOLIC, center=(cx, cy),
                     radius=float(info["radius"])),
           P.make_bc(P.BC_PRESSURE, 2, -1, rho=1.0)]
    dom = P.Domain(ctx, dims, flags, 0.7, bcs)
    dom.step(101)
    rho, u = dom.macros()
    fl = flags > 0
    vtu = os.path.join(out, "Fluid_Variables/100_rank_0.vtu")
    assert np.array_equal(vtu_array(vtu, "Density"), rho[fl])
    assert np.array_equal(vtu_array(vtu, "Velocity").reshape(-1, 3), u[:, fl].T)
    wss = dom.wss()
    w_vtu = vtu_array(os.path.join(out, "Wall_Shear_Stress_VTK/100_rank_0_wall_shear_stress.vtu"), "WallShearStress")
    walls = np.nonzero(flags == 2)[0]
    nx, ny, nz = dims
    k, rem = walls // (nx * ny), walls % (nx * ny)
    order = np.lexsort((k, rem // nx, rem % nx))  # i outermost, then j, then k
    assert np.array_equal(w_vtu, wss[walls[order]])


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
