"""A small bent-pipe case written as STL + configuration for lbm_run."""
import os

from pv2013_amd.geometry import bent_pipe, write_stl_ascii


def write_case(dirpath, steps=101, every=100, storage="f64", length=40, radius=6, bend=None):
    tri, info = bent_pipe(length, radius, bend)
    stl = os.path.join(dirpath, "pipe.stl")
    write_stl_ascii(stl, tri)
    nx, ny, nz = info["dims"]
    cx, cy = info["inlet_center"]
    cfg = os.path.join(dirpath, "case.cfg")
    with open(cfg, "w") as fh:
        fh.write("# bent pipe\nstl_file %s\nspacing 1\nbox_origin 0 0 0\nbox_dims %d %d %d\n" % (stl, nx, ny, nz))
        fh.write("plane inlet z 1\nplane outlet z %d\n" % info["outlet_z"])
        fh.write("bc inlet velocity 1 0 0 0.03 parabolic %r %r %r\n" % (cx, cy, float(radius)))
        fh.write("bc outlet pressure -1 1.0\n")
        fh.write("lattice 19\nstorage %s\nquirks 1\ntau 0.7\nsteps %d\noutput_every %d\noutput_dir %s\n"
                 % (storage, steps, every, os.path.join(dirpath, "out")))
    return cfg, stl, info
