"""Flag boxes and boundary-condition sets shared by the CPU and GPU tests."""
import numpy as np


def pipe_flags(dims, radius, center=None, axis=2):
    """Straight circular pipe along `axis` (inside: r < radius), flags x-fastest."""
    nx, ny, nz = dims
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    coords = [i, j, k]
    t = [a for a in range(3) if a != axis]
    if center is None:
        center = ((dims[t[0]] - 1) / 2.0, (dims[t[1]] - 1) / 2.0)
    r2 = (coords[t[0]] - center[0]) ** 2 + (coords[t[1]] - center[1]) ** 2
    return (r2 < radius * radius).astype(np.uint8).reshape(-1), center


def box_flags(dims, lo, hi):
    nx, ny, nz = dims
    f = np.zeros((nz, ny, nx), np.uint8)
    f[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] = 1
    return f.reshape(-1)


def pipe_case(oracle, dims, radius, u_max, rho_out=1.0, axis=2, waveform=-1, typo=0):
    """Pipe with a parabolic velocity inlet on the low face and a pressure
    outlet on the high face of `axis`; returns (flags, bcs as dicts)."""
    flags, center = pipe_flags(dims, radius, axis=axis)
    n = dims[axis]
    planes = [oracle.plane(axis, "0"), oracle.plane(axis, str(n - 1))]
    g = oracle.make_grid((0, 0, 0), 1.0, dims, (0, dims[0] - 1, 0, dims[1] - 1, 0, dims[2] - 1))
    cl = oracle.classify(g, flags, planes)
    u = [0.0, 0.0, 0.0]
    u[axis] = u_max
    bcs = [dict(kind=0, axis=axis, sigma=1, u=tuple(u), profile=1, center=center, radius=radius,
                waveform=waveform),
           dict(kind=1, axis=axis, sigma=-1, rho=rho_out, typo=typo)]
    return cl, bcs
