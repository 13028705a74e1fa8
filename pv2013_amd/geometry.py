"""Seeded synthetic vessel meshes for the benchmark cases.

Surfaces are extracted from an implicit vessel (union of tubes, capped by two
planes) with marching tetrahedra, so every mesh is closed and the +z ray
parity of the voxelizer is well defined.  The extraction grid is offset by
irrational fractions of a lattice cell so that no lattice column runs along a
cap edge (the reference's inclusive edge test counts such edges twice).

All coordinates are in lattice units (spacing 1, lattice point (0,0,0) at the
origin).  Triangles are (n, 12) float64: v1, v2, v3, unit normal (outward).
"""
import math

import numpy as np

_OFFSET = np.array([0.1234567, 0.3456789, 0.0])

# Kuhn decomposition of the unit cube into 6 tetrahedra along the 0-7 diagonal
_CORNER = np.array([[(c >> a) & 1 for a in range(3)] for c in range(8)])
_TETS = []
for _p in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
    v = [0]
    acc = 0
    for a in _p:
        acc |= 1 << a
        v.append(acc)
    _TETS.append(v)


def _case_tables():
    """For each of the 16 inside/outside patterns of a tetrahedron, the
    triangles as triples of edges (i, j) between an inside and outside vertex."""
    table = {}
    for case in range(1, 15):
        ins = [i for i in range(4) if (case >> i) & 1]
        out = [i for i in range(4) if not (case >> i) & 1]
        if len(ins) == 1:
            table[case] = [[(ins[0], out[0]), (ins[0], out[1]), (ins[0], out[2])]]
        elif len(ins) == 3:
            table[case] = [[(ins[0], out[0]), (ins[1], out[0]), (ins[2], out[0])]]
        else:
            a, b = ins
            c, d = out
            table[case] = [[(a, c), (a, d), (b, d)], [(a, c), (b, d), (b, c)]]
    return table


_TABLE = _case_tables()


def marching_tets(sdf, lo, hi, h, slab=8):
    """Zero level set of sdf (negative inside) on [lo, hi] with cell size h."""
    lo = np.asarray(lo, float) + _OFFSET * h
    n = np.maximum(np.ceil((np.asarray(hi, float) - lo) / h).astype(int), 1)
    xs = lo[0] + h * np.arange(n[0] + 1)
    ys = lo[1] + h * np.arange(n[1] + 1)
    out = []
    for k0 in range(0, n[2], slab):
        k1 = min(n[2], k0 + slab)
        zs = lo[2] + h * np.arange(k0, k1 + 1)
        X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
        S = sdf(X, Y, Z)
        if not ((S < 0).any() and (S >= 0).any()):
            continue
        nx, ny, nz = S.shape[0] - 1, S.shape[1] - 1, S.shape[2] - 1
        gi, gj, gk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
        gi, gj, gk = gi.ravel(), gj.ravel(), gk.ravel()
        cv = S[:-1, :-1, :-1].ravel()
        sel = np.zeros(cv.shape, bool)
        cvals = []
        for c in range(8):
            dx, dy, dz = _CORNER[c]
            vals = S[dx:dx + nx, dy:dy + ny, dz:dz + nz].ravel()
            cvals.append(vals)
            sel |= (vals < 0) != (cv < 0)
        idx = np.nonzero(sel)[0]
        if idx.size == 0:
            continue
        gi, gj, gk = gi[idx], gj[idx], gk[idx]
        cvals = [v[idx] for v in cvals]
        cpos = [np.stack([xs[gi + _CORNER[c][0]], ys[gj + _CORNER[c][1]], zs[gk + _CORNER[c][2]]], 1)
                for c in range(8)]
        for tet in _TETS:
            v = np.stack([cvals[t] for t in tet], 1)
            p = np.stack([cpos[t] for t in tet], 1)
            case = ((v < 0) * (1 << np.arange(4))).sum(1)
            for cs, tris in _TABLE.items():
                m = case == cs
                if not m.any():
                    continue
                vv, pp = v[m], p[m]
                ins = (cs >> np.arange(4)) & 1
                pin = pp[:, ins == 1].mean(1)
                pout = pp[:, ins == 0].mean(1)
                for tri in tris:
                    verts = []
                    for (i, j) in tri:
                        t = vv[:, i] / (vv[:, i] - vv[:, j])
                        verts.append(pp[:, i] + t[:, None] * (pp[:, j] - pp[:, i]))
                    a, b, c = verts
                    nrm = np.cross(b - a, c - a)
                    flip = (nrm * (pout - pin)).sum(1) < 0
                    b2 = np.where(flip[:, None], c, b)
                    c2 = np.where(flip[:, None], b, c)
                    out.append(np.concatenate([a, b2, c2], 1))
    if not out:
        return np.zeros((0, 12))
    V = np.concatenate(out, 0)
    nrm = np.cross(V[:, 3:6] - V[:, 0:3], V[:, 6:9] - V[:, 0:3])
    ln = np.linalg.norm(nrm, axis=1)
    keep = ln > 1e-12
    V, nrm, ln = V[keep], nrm[keep], ln[keep]
    return np.concatenate([V, nrm / ln[:, None]], 1)


class WallModes:
    """Radial wall perturbation: sum of seeded Fourier modes in (angle, axial
    position) whose amplitudes add up to `amp` lattice units."""

    def __init__(self, rng, nmodes=8, amp=1.5, length=512.0):
        self.m = rng.integers(1, 7, nmodes)
        self.q = 2 * np.pi * rng.integers(1, 5, nmodes) / length
        self.phi = 2 * np.pi * rng.random(nmodes)
        a = rng.random(nmodes) + 0.2
        self.a = amp * a / a.sum()

    def __call__(self, theta, s):
        out = np.zeros(np.broadcast(theta, s).shape)
        for k in range(len(self.a)):
            out += self.a[k] * np.cos(self.m[k] * theta + self.q[k] * s + self.phi[k])
        return out


def _tube_dist(dx, dy, r0, modes, s, cosang=1.0):
    rh = np.hypot(dx, dy)
    theta = np.arctan2(dy, dx)
    return rh * cosang - (r0 + (modes(theta, s) if modes is not None else 0.0))


def y_bifurcation(seed, dims=(256, 256, 512), parent_r=40.0, daughter_r=30.0, angle_deg=25.0,
                  split_z=256.0, amp=1.5, nmodes=8, h=3.1, spread=80.0):
    """Parent tube along z that splits at split_z into two daughters leaving at
    +-angle in the x-z plane.  The daughters bend back towards z so they stay
    inside the box: their centreline offset is spread*(1 - exp(-tan(a) dz/spread)).
    Capped at z = 0.5 and z = nz - 1.5.  Returns (triangles, info)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    cx, cy = (nx - 1) / 2.0, (ny - 1) / 2.0
    ta = math.tan(math.radians(angle_deg))
    mp = WallModes(rng, nmodes, amp, nz)
    md = [WallModes(rng, nmodes, amp, nz), WallModes(rng, nmodes, amp, nz)]
    z0, z1 = 0.5, nz - 1.5

    def sdf(X, Y, Z):
        d = _tube_dist(X - cx, Y - cy, parent_r, mp, Z)
        d = np.maximum(d, Z - (split_z + parent_r))
        for sgn, modes in ((1.0, md[0]), (-1.0, md[1])):
            dz = np.maximum(Z - split_z, 0.0)
            off = spread * (1.0 - np.exp(-ta * dz / spread))
            slope = ta * np.exp(-ta * dz / spread)
            cosang = 1.0 / np.sqrt(1.0 + slope * slope)
            dd = _tube_dist(X - (cx + sgn * off), Y - cy, daughter_r, modes, Z, cosang)
            dd = np.maximum(dd, (split_z - daughter_r) - Z)
            d = np.minimum(d, dd)
        return np.maximum(d, np.maximum(z0 - Z, Z - z1))

    tri = marching_tets(sdf, (cx - 130, cy - 50, 0.0), (cx + 130, cy + 50, nz - 1.0), h)
    info = {"dims": dims, "center": (cx, cy), "inlet_z": 1, "outlet_z": nz - 2,
            "parent_r": parent_r, "daughter_offset": spread * (1 - math.exp(-ta * (z1 - split_z) / spread))}
    return tri, info


def bent_pipe(length, radius, curvature_radius=None, h=None, pad=3):
    """Pipe along z of the given length whose centreline bends in x as
    x_c(z) = A (1 - cos(2 pi z / L)) with peak curvature 1/curvature_radius
    (A = L^2 / (4 pi^2 R_c)); both ends are straight and axis-aligned so the
    inlet (z = 1) and outlet (z = nz - 2) are lattice planes.  No bend when
    curvature_radius is None.  Returns (triangles, info) in lattice units."""
    L = float(length)
    amp = 0.0 if not curvature_radius else L * L / (4 * np.pi ** 2 * curvature_radius)
    nx = int(math.ceil(2 * radius + 2 * amp)) + 2 * pad
    ny = int(math.ceil(2 * radius)) + 2 * pad
    nz = int(length) + 2
    cx0 = pad + radius
    cy = (ny - 1) / 2.0
    z0, z1 = 0.5, nz - 1.5
    k = 2 * np.pi / (z1 - z0)

    def sdf(X, Y, Z):
        xc = cx0 + amp * (1 - np.cos(k * (Z - z0)))
        slope = amp * k * np.sin(k * (Z - z0))
        cosang = 1.0 / np.sqrt(1.0 + slope * slope)
        d = _tube_dist(X - xc, Y - cy, radius, None, Z, cosang)
        return np.maximum(d, np.maximum(z0 - Z, Z - z1))

    if h is None:
        h = min(1.0, max(0.5, radius / 8.0))
    tri = marching_tets(sdf, (0.0, 0.0, 0.0), (nx - 1.0, ny - 1.0, nz - 1.0), h)
    info = {"dims": (nx, ny, nz), "inlet_center": (cx0, cy), "radius": radius, "inlet_z": 1,
            "outlet_z": nz - 2, "bend": amp}
    return tri, info


def _catmull_rom(P, samples_per_seg=200):
    """Centripetal-free uniform Catmull-Rom spline through control points P."""
    P = np.asarray(P, float)
    ext = np.vstack([2 * P[0] - P[1], P, 2 * P[-1] - P[-2]])
    out = []
    t = np.linspace(0, 1, samples_per_seg, endpoint=False)[:, None]
    for i in range(1, len(ext) - 2):
        p0, p1, p2, p3 = ext[i - 1], ext[i], ext[i + 1], ext[i + 2]
        out.append(0.5 * ((2 * p1) + (-p0 + p2) * t + (2 * p0 - 5 * p1 + 4 * p2 - p3) * t * t
                          + (-p0 + 3 * p1 - 3 * p2 + p3) * t * t * t))
    out.append(P[-1][None, :])
    return np.vstack(out)


def tortuous_tube(seed, n=512, r_min=10.0, r_max=60.0, ncontrol=7, h=2.2):
    """Closed tube around a seeded random cubic (Catmull-Rom) centreline inside
    an n^3 box, radius varying smoothly in [r_min, r_max]; capped where the
    centreline ends.  About 500k facets at the default h for n = 512."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    m = r_max + 8.0
    ctrl = m + (n - 1 - 2 * m) * rng.random((ncontrol, 3))
    ctrl[0, 2], ctrl[-1, 2] = m, n - 1 - m
    ctrl = ctrl[np.argsort(ctrl[:, 2])]
    cl = _catmull_rom(ctrl)
    cl = np.clip(cl, m, n - 1 - m)
    seg = np.linalg.norm(np.diff(cl, axis=0), axis=1)
    s = np.concatenate([[0.0], np.cumsum(seg)])
    smax = s[-1]
    # radius along the arc length: smooth random mix of a few cosines
    ks = rng.integers(1, 5, 3)
    ph = 2 * np.pi * rng.random(3)
    wave = sum(np.cos(2 * np.pi * k * s / smax + p) for k, p in zip(ks, ph)) / 3.0
    rad = r_min + (r_max - r_min) * 0.5 * (1 + wave)
    # resample densely so the nearest sample approximates the curve distance
    tree = cKDTree(cl)
    tangent_end = [cl[1] - cl[0], cl[-1] - cl[-2]]
    tangent_end = [t / np.linalg.norm(t) for t in tangent_end]

    def sdf(X, Y, Z):
        pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
        dist, idx = tree.query(pts)
        d = dist - rad[idx]
        # flat caps at both ends of the centreline
        c0 = ((pts - cl[0]) * tangent_end[0]).sum(1)
        c1 = ((pts - cl[-1]) * tangent_end[1]).sum(1)
        d = np.maximum(d, np.maximum(-c0, c1))
        return d.reshape(X.shape)

    tri = marching_tets(sdf, (0.0, 0.0, 0.0), (n - 1.0, n - 1.0, n - 1.0), h, slab=4)
    return tri, {"dims": (n, n, n), "centreline_length": smax}


def vessel_tree(seed, dims=(768, 384, 384), r_root=60.0, r_leaf=15.0, levels=3, amp=1.5, nmodes=8, h=3.0):
    """Seeded multi-branch tree along +x: a root tube from the inlet plane
    x = 0.5 that bifurcates `levels` times (alternately in y and z), radii
    shrinking geometrically from r_root to r_leaf, every leaf running parallel
    to x to the outlet plane x = nx - 1.5; wall perturbation as in the
    Y-bifurcation.  Returns (triangles, info)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    cy, cz = (ny - 1) / 2.0, (nz - 1) / 2.0
    x0, x1 = 0.5, nx - 1.5
    seglen = (x1 - x0) / (levels + 1)
    ratio = (r_leaf / r_root) ** (1.0 / levels)
    # branch k at level l: lateral offset of its centre, reached with a smooth step
    spread_y, spread_z = (ny - 2 * r_root) / 4.0, (nz - 2 * r_root) / 4.0
    frontier = [(cy, cz)]
    tubes = []
    for lev in range(levels + 1):
        r = r_root * ratio ** lev
        xs = x0 + lev * seglen
        nxt = []
        for (py, pz) in frontier:
            tubes.append((xs, r, py, pz))
            if lev < levels:
                dy = spread_y / (2 ** (lev // 2)) if lev % 2 == 0 else 0.0
                dz = spread_z / (2 ** (lev // 2)) if lev % 2 == 1 else 0.0
                nxt += [(py - dy, pz - dz), (py + dy, pz + dz)]
        frontier = nxt
    modes = [WallModes(rng, nmodes, amp, nx) for _ in tubes]
    # each tube: from x = xs (where it leaves its parent's end position) to x1,
    # its centre moving smoothly from the parent's centre to its own
    tube_par = []
    lev_of = []
    k = 0
    for lev in range(levels + 1):
        for j in range(2 ** lev):
            xs, r, py, pz = tubes[k]
            if lev == 0:
                tube_par.append((cy, cz))
            else:
                tube_par.append(tubes[(2 ** (lev - 1) - 1) + j // 2][2:4])
            lev_of.append(lev)
            k += 1

    def sdf(X, Y, Z):
        d = np.full(X.shape, np.inf)
        for (xs, r, py, pz), (qy, qz), md, lev in zip(tubes, tube_par, modes, lev_of):
            t = np.clip((X - xs) / (0.6 * seglen), 0.0, 1.0)
            sm = t * t * (3 - 2 * t)
            yc = qy + (py - qy) * sm
            zc = qz + (pz - qz) * sm
            dt = _tube_dist(Y - yc, Z - zc, r, md, X)
            # flat tube ends sit off lattice planes (irrational fractions): a
            # voxelizer ray running inside a cap's plane would be degenerate
            dt = np.maximum(dt, (xs - r * 0.5 - 0.2718281828) - X)
            if lev < levels:  # a parent ends inside its children's start
                dt = np.maximum(dt, X - (xs + seglen + 0.5 * r + 0.3141592654))
            d = np.minimum(d, dt)
        return np.maximum(d, np.maximum(x0 - X, X - x1))

    tri = marching_tets(sdf, (0.0, 0.0, 0.0), (nx - 1.0, ny - 1.0, nz - 1.0), h, slab=6)
    info = {"dims": dims, "inlet_x": 1, "outlet_x": nx - 2, "root_center": (cy, cz), "r_root": r_root,
            "leaves": 2 ** levels}
    return tri, info


def fourier_waveform(seed, nk=4, amp=0.5):
    """Seeded a_k, phi_k for u(t) = U (1 + sum a_k sin(2 pi k t / T + phi_k))."""
    rng = np.random.default_rng(seed + 7919)
    a = rng.random(nk)
    a = amp * a / a.sum() if a.sum() > 0 else a
    phi = 2 * np.pi * rng.random(nk)
    return a, phi


def write_stl_ascii(path, tri, name="vessel", exact=False):
    """ASCII STL in the layout Reader::readSTLASCII parses (GeometryReader.cpp:
    33-117).  exact: 17 significant digits, so every double reads back
    unchanged; otherwise 10."""
    f = "%.17g" if exact else "%.9e"
    row = ("facet normal {0} {0} {0}\n outer loop\n  vertex {0} {0} {0}\n  vertex {0} {0} {0}\n"
           "  vertex {0} {0} {0}\n endloop\nendfacet\n").format(f)
    with open(path, "w") as fh:
        fh.write("solid %s\n" % name)
        fh.write("".join(row % (t[9], t[10], t[11], *t[:9]) for t in tri))
        fh.write("endsolid %s\n" % name)


def write_stl_binary(path, tri, name="vessel"):
    """Binary STL (the layout Reader::readSTLBinary reads, GeometryReader.cpp:
    119-176): an 80-byte header, a uint32 facet count, then per facet the normal
    and three vertices as little-endian float32 and a zero uint16.  Values are
    rounded to float32."""
    tri = np.asarray(tri, np.float64)
    rec = np.zeros(len(tri), dtype=[("n", "<f4", 3), ("v", "<f4", 9), ("attr", "<u2")])
    rec["n"] = tri[:, 9:12]
    rec["v"] = tri[:, :9]
    with open(path, "wb") as fh:
        fh.write(("binary STL " + name).encode()[:80].ljust(80, b" "))
        fh.write(np.uint32(len(tri)).tobytes())
        fh.write(rec.tobytes())
