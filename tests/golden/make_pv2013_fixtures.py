"""Write tests/golden/pv2013.npz from the data files the reference ships.

Inputs (data, not code), read from the reference checkout:
  PV2013.stl                               the vessel mesh (4988 facets)
  InboundaryPointsFromGeometry.vtk         inlet boundary points
  OutboundaryPointsFromGeometry{1..6}.vtk  outlet boundary points
  WallboundaryPointsFromGeometry.vtk       wall boundary points

The VTK files were written by Writer::writePointsToFile
(GeometryWriter.cpp:35-68) with the default 6-significant-digit stream
precision; each point is mapped back to its grid index
round((p - bbox_min) / spacing) with spacing 0.1 (Input.h:9), which is
unambiguous (the largest fractional error is far below half a cell).

Output arrays:
  tri        (n, 12) float64: v1, v2, v3, file normal, parsed as the reference's
             readSTLASCII does (GeometryReader.cpp:33-117)
  <set>      sorted int64 site indices i + nx*(j + ny*k) for set in
             inlet, out1..out6, wall
  dims       grid dims implied by the bbox (ComplexGeometry.cpp:68-72)

and tests/golden/pv2013_points_sha256.json: the SHA-256 digest of each of the
eight boundary-point files, so that files written by the engine or the oracle
can be checked for byte identity with the reference's.

Run:  python tests/golden/make_pv2013_fixtures.py [/root/reference]
"""
import hashlib
import json
import math
import os
import sys

import numpy as np

SPACING = 0.1


def parse_stl(path):
    tris = []
    cur = [0.0] * 12
    with open(path) as fh:
        lines = fh.read().splitlines()
    it = iter(lines)
    for line in it:
        tok = line.split()
        if len(tok) < 5 or tok[0] != "facet" or tok[1] != "normal":
            continue
        cur[9:12] = [float(t) for t in tok[2:5]]
        tok = next(it).split()
        if len(tok) < 2 or tok[0] != "outer" or tok[1] != "loop":
            continue
        for v in range(3):
            tok = next(it).split()
            if len(tok) < 4 or tok[0] != "vertex":
                break
            cur[3 * v:3 * v + 3] = [float(t) for t in tok[1:4]]
        tok = next(it).split()
        if not tok or tok[0] != "endloop":
            continue
        tok = next(it).split()
        if not tok or tok[0] != "endfacet":
            continue
        tris.append(list(cur))
    return np.array(tris, dtype=np.float64)


def parse_vtk_points(path):
    with open(path) as fh:
        lines = fh.read().splitlines()
    for n, line in enumerate(lines):
        if line.startswith("POINTS"):
            count = int(line.split()[1])
            pts = np.array([[float(v) for v in lines[n + 1 + i].split()] for i in range(count)])
            return pts
    raise ValueError(path)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    tri = parse_stl(os.path.join(ref, "PV2013.stl"))
    v = tri[:, :9].reshape(-1, 3, 3)
    bmin = v.min(axis=(0, 1))
    bmax = v.max(axis=(0, 1))
    dims = [int(math.ceil((bmax[a] - bmin[a]) / SPACING)) + 1 for a in range(3)]
    out = {"tri": tri, "dims": np.array(dims, dtype=np.int64)}
    names = {"inlet": "InboundaryPointsFromGeometry.vtk",
             "wall": "WallboundaryPointsFromGeometry.vtk"}
    for k in range(1, 7):
        names["out%d" % k] = "OutboundaryPointsFromGeometry%d.vtk" % k
    worst = 0.0
    for key, fname in names.items():
        pts = parse_vtk_points(os.path.join(ref, fname))
        g = (pts - bmin) / SPACING
        idx = np.rint(g).astype(np.int64)
        worst = max(worst, float(np.abs(g - idx).max()) if len(g) else 0.0)
        site = idx[:, 0] + dims[0] * (idx[:, 1] + dims[1] * idx[:, 2])
        out[key] = np.sort(site)
        print(key, len(site))
    print("max fractional index error", worst)
    digests = {fname: hashlib.sha256(open(os.path.join(ref, fname), "rb").read()).hexdigest()
               for fname in names.values()}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "pv2013_points_sha256.json"), "w") as fh:
        json.dump(digests, fh, indent=1, sort_keys=True)
        fh.write("\n")
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pv2013.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, os.path.getsize(dst), "bytes")


if __name__ == "__main__":
    main()
