"""The C++ host driver (lbm_run, the reference's main over the C ABI): its
configuration parser fails cleanly without touching a GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "pv2013_amd", "csrc", "lbm_run")


@pytest.fixture(scope="module")
def run():
    if not os.path.exists(RUN):
        subprocess.check_call(["make", "-s", "-C", os.path.dirname(RUN), "lbm_run"])
    return RUN


def test_usage_and_config_errors(run, tmp_path):
    assert subprocess.run([run], capture_output=True).returncode == 2
    bad = tmp_path / "bad.cfg"
    bad.write_text("stl_file x.stl\nplane inlet q 1\n")
    r = subprocess.run([run, str(bad)], capture_output=True, text=True)
    assert r.returncode == 2 and "bad plane line" in r.stderr
    nobc = tmp_path / "nobc.cfg"
    nobc.write_text("stl_file x.stl\nplane inlet z 1\n")
    r = subprocess.run([run, str(nobc)], capture_output=True, text=True)
    assert r.returncode == 2 and "has no bc line" in r.stderr
    unk = tmp_path / "unk.cfg"
    unk.write_text("stl_file x.stl\nplane inlet z 1\nbc outlet pressure -1 1.0\n")
    r = subprocess.run([run, str(unk)], capture_output=True, text=True)
    assert r.returncode == 2 and "unknown plane" in r.stderr


def test_valid_config_reaches_the_device(run, tmp_path):
    from host_case import write_case
    cfg, _, _ = write_case(str(tmp_path))
    r = subprocess.run([run, cfg], capture_output=True, text=True, timeout=120)
    # 0 with a GPU; without one the parse succeeded and the device open failed
    assert r.returncode in (0, 1), r.stderr
    if r.returncode == 1:
        assert "no GPU device" in r.stderr
