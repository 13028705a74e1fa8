"""The benchmark configurations, built through the product path (seeded STL ->
GPU voxelizer -> GPU classification -> boundary conditions).

  c1  D3Q19 FP64 straight pipe, 32x32x128, R 12, tau 0.8, parabolic inlet
      u_max 0.05, outlet rho 1.0, 2000 steps
  c2  D3Q19 256x256x512 Y-bifurcation, pulsatile parabolic inlet, 8000 steps in
      FP64 / FP32 / FP16 shifted storage
  c3  voxelizer only: tortuous tube (~500k facets) in a 512^3 grid
  c4  64 straight and bent pipes (lengths 64-512, radii 8-40, curvature radius
      100-1000), FP32, each with its own pulsatile inlet, packed into one domain,
      4000 steps
  c5  D3Q27 768x384x384 vessel tree, FP16 shifted storage, 10000 steps; WSS
      against an FP64 run of 2000 steps
"""
import numpy as np

from . import capi as P
from .geometry import (bent_pipe, fourier_waveform, tortuous_tube, vessel_tree, y_bifurcation)


def lattice_grid(dims):
    return P.make_grid((0.0, 0.0, 0.0), 1.0, dims,
                       (0.0, dims[0] - 1.0, 0.0, dims[1] - 1.0, 0.0, dims[2] - 1.0))


def pipe_flags(dims, radius):
    nx, ny, nz = dims
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    cx, cy = (nx - 1) / 2.0, (ny - 1) / 2.0
    return (((i - cx) ** 2 + (j - cy) ** 2) < radius * radius).astype(np.uint8).reshape(-1), (cx, cy)


def c1(ctx, dims=(32, 32, 128), radius=12.0, u_max=0.05):
    flags, (cx, cy) = pipe_flags(dims, radius)
    flags, _ = ctx.classify(lattice_grid(dims), flags, [P.make_plane(2, "0"), P.make_plane(2, str(dims[2] - 1))])
    bcs = [dict(kind=P.BC_VELOCITY, axis=2, sigma=1, u=(0.0, 0.0, u_max), profile=P.PROFILE_PARABOLIC,
                center=(cx, cy), radius=radius),
           dict(kind=P.BC_PRESSURE, axis=2, sigma=-1, rho=1.0)]
    return {"name": "c1", "dims": dims, "flags": flags, "bcs": bcs, "tau": 0.8, "waves": {}, "steps": 2000}


def c2(ctx, seed=2013, dims=(256, 256, 512), tau=0.66, period=4000.0):
    tri, info = y_bifurcation(seed, dims)
    g = lattice_grid(dims)
    flags, vox_ms = ctx.voxelize(tri, g)
    cx, cy = info["center"]
    planes = [P.make_plane(2, "1"),
              P.make_plane(2, str(info["outlet_z"]), ranges=((0, -1.0, cx),)),
              P.make_plane(2, str(info["outlet_z"]), ranges=((0, cx, dims[0] + 1.0),))]
    flags, cls_ms = ctx.classify(g, flags, planes)
    bcs = [dict(kind=P.BC_VELOCITY, axis=2, sigma=1, u=(0.0, 0.0, 0.03), profile=P.PROFILE_PARABOLIC,
                center=(cx, cy), radius=info["parent_r"], waveform=0),
           dict(kind=P.BC_PRESSURE, axis=2, sigma=-1, rho=1.0),
           dict(kind=P.BC_PRESSURE, axis=2, sigma=-1, rho=1.0)]
    a, phi = fourier_waveform(seed)
    return {"name": "c2", "dims": dims, "flags": flags, "bcs": bcs, "tau": tau,
            "waves": {0: (P.WAVE_FOURIER, period, a, phi)}, "steps": 8000, "ntri": len(tri),
            "vox_ms": vox_ms, "cls_ms": cls_ms}


def c3_mesh(seed=3, n=512, h=2.2):
    tri, info = tortuous_tube(seed, n=n, h=h)
    return tri, lattice_grid((n, n, n))


def c4(ctx, seed=4, ncases=64, tau=0.7, period=2000.0):
    rng = np.random.default_rng(seed)
    cases = []
    for c in range(ncases):
        length = int(rng.integers(64, 513))
        radius = float(rng.integers(8, 41))
        bend = None if c % 2 == 0 else float(rng.uniform(100.0, 1000.0))
        tri, info = bent_pipe(length, radius, bend)
        dims = info["dims"]
        g = lattice_grid(dims)
        flags, _ = ctx.voxelize(tri, g)
        flags, _ = ctx.classify(g, flags, [P.make_plane(2, "1"), P.make_plane(2, str(info["outlet_z"]))])
        cx, cy = info["inlet_center"]
        bcs = [dict(kind=P.BC_VELOCITY, axis=2, sigma=1, u=(0.0, 0.0, 0.02), profile=P.PROFILE_PARABOLIC,
                    center=(cx, cy), radius=radius, waveform=0),
               dict(kind=P.BC_PRESSURE, axis=2, sigma=-1, rho=1.0)]
        a, phi = fourier_waveform(seed * 1000 + c)
        cases.append({"name": "c4_%02d" % c, "dims": dims, "flags": flags, "bcs": bcs, "tau": tau,
                      "waves": {0: (P.WAVE_FOURIER, period, a, phi)}, "length": length, "radius": radius,
                      "bend": bend})
    return cases


def c5(ctx, seed=5, dims=(768, 384, 384), tau=0.6, h=3.0):
    tri, info = vessel_tree(seed, dims, h=h)
    g = lattice_grid(dims)
    flags, vox_ms = ctx.voxelize(tri, g)
    cy, cz = info["root_center"]
    flags, _ = ctx.classify(g, flags, [P.make_plane(0, "1"), P.make_plane(0, str(info["outlet_x"]))])
    bcs = [dict(kind=P.BC_VELOCITY, axis=0, sigma=1, u=(0.03, 0.0, 0.0), profile=P.PROFILE_PARABOLIC,
                center=(cy, cz), radius=info["r_root"]),
           dict(kind=P.BC_PRESSURE, axis=0, sigma=-1, rho=1.0)]
    return {"name": "c5", "dims": dims, "flags": flags, "bcs": bcs, "tau": tau, "waves": {}, "steps": 10000,
            "ntri": len(tri), "vox_ms": vox_ms}


def make_domain(ctx, case, q=19, storage=P.STORE_F32, **kw):
    dom = P.Domain(ctx, case["dims"], case["flags"], case["tau"], [P.make_bc(**b) for b in case["bcs"]],
                   q=q, storage=storage, **kw)
    for slot, (kind, period, a, phi) in case["waves"].items():
        dom.set_waveform(slot, kind, period, a, phi)
    return dom


def errors(ref, got, mask=None):
    """L2 (relative) and L-infinity (relative to max |ref|) errors."""
    ref = np.asarray(ref)
    got = np.asarray(got)
    if mask is not None:
        ref = ref[..., mask]
        got = got[..., mask]
    den = np.linalg.norm(ref)
    mx = np.abs(ref).max()
    return {"l2": float(np.linalg.norm(got - ref) / den) if den > 0 else 0.0,
            "linf": float(np.abs(got - ref).max() / mx) if mx > 0 else 0.0}
