"""Reader for the VTK XML unstructured-grid files lbm_run and orc_run write:
inline ASCII data arrays or one appended raw block (UInt64 byte counts)."""
import re

import numpy as np

_DT = {"Float64": np.float64, "Int64": np.int64, "UInt8": np.uint8}


def read_vtu(path):
    """{name: array} for every DataArray; the points are under "Points"
    (shape (n, 3)), multi-component arrays are (n, ncomp)."""
    data = open(path, "rb").read()
    cut = data.find(b"<AppendedData")
    head = data[: cut if cut >= 0 else len(data)].decode()
    block = None
    if cut >= 0:
        block = data.index(b"_", cut) + 1
    out = {}
    in_points = False
    for m in re.finditer(r"<Points>|</Points>|<DataArray([^>]*?)(/?)>", head):
        tag = m.group(0)
        if tag == "<Points>":
            in_points = True
            continue
        if tag == "</Points>":
            in_points = False
            continue
        attrs = dict(re.findall(r'(\w+)="([^"]*)"', m.group(1)))
        name = "Points" if in_points else attrs["Name"]
        dt = _DT[attrs["type"]]
        ncomp = int(attrs.get("NumberOfComponents", "1"))
        if attrs["format"] == "appended":
            off = block + int(attrs["offset"])
            nbytes = int(np.frombuffer(data, np.uint64, 1, off)[0])
            arr = np.frombuffer(data, dt, nbytes // np.dtype(dt).itemsize, off + 8).copy()
        else:
            end = head.index("</DataArray>", m.end())
            arr = np.array(head[m.end():end].split(), dtype=dt)
        out[name] = arr.reshape(-1, ncomp) if ncomp > 1 else arr
    return out


def read_points_vtk(path):
    """Points of a legacy VTK point file (Writer::writePointsToFile format)."""
    lines = open(path).read().splitlines()
    for n, line in enumerate(lines):
        if line.startswith("POINTS"):
            count = int(line.split()[1])
            return np.array([[float(v) for v in lines[n + 1 + i].split()] for i in range(count)]).reshape(-1, 3)
    raise ValueError(path)
