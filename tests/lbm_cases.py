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


def poiseuille_metrics(dims, radius, tau, flags, rho, u, axis=2):
    """Analytic checks of developed pipe flow along z (x, y centred).

    Returns, at cross-sections z = n/4, n/2, 3n/4: the centreline velocity u_c and
    effective radius R_eff of a least-squares parabola u_c (1 - r^2 / R_eff^2)
    fitted to the nodes with r < R - 1, its relative residual, and the
    Hagen-Poiseuille prediction u_c = G R_eff^2 / (4 rho nu) from the measured
    pressure gradient G = -(1/3) d<rho>/dz (p = rho c_s^2, nu = (tau - 1/2)/3);
    plus the spread of the mass flux sum(rho u_z) along the pipe."""
    assert axis == 2
    nx, ny, nz = dims
    uz = u[2].reshape(nz, ny, nx)
    rr = rho.reshape(nz, ny, nx)
    fl = flags.reshape(nz, ny, nx) > 0
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    r2 = ((i - (nx - 1) / 2) ** 2 + (j - (ny - 1) / 2) ** 2).astype(float)
    inner = r2 < (radius - 1.0) ** 2
    A = np.stack([np.ones(inner.sum()), r2[inner]], 1)
    rm = np.array([rr[k][fl[k]].mean() for k in range(nz)])
    ks = np.arange(nz // 6, nz - nz // 6)
    G = -np.polyfit(ks, rm[ks], 1)[0] / 3.0
    nu = (tau - 0.5) / 3.0
    out = {"sections": []}
    for k in (nz // 4, nz // 2, 3 * nz // 4):
        prof = uz[k][inner]
        (a, b), *_ = np.linalg.lstsq(A, prof, rcond=None)
        reff2 = -a / b
        out["sections"].append({
            "z": k, "u_c": a, "R_eff": float(np.sqrt(reff2)),
            "residual": float(np.linalg.norm(A @ [a, b] - prof) / np.linalg.norm(prof)),
            "u_c_hagen_poiseuille": G * reff2 / (4 * nu * rm[k])})
    flux = np.array([(rr[k] * uz[k]).sum() for k in range(1, nz - 1)])
    out["flux_spread"] = float((flux.max() - flux.min()) / flux.mean())
    out["G"] = G
    return out


def assert_poiseuille(m, radius, u_max):
    """Bounds for config 1 (32x32x128, R 12, tau 0.8, u_max 0.05, 2000 steps,
    clean bounce-back): the profile is a parabola to 2%, its radius within one
    lattice unit of R, its centreline within 10% of the inlet's u_max (the
    density falls along the pipe, so rho u is what stays constant), the
    Hagen-Poiseuille relation holds to 3% and the mass flux is the same at
    every section to 1%."""
    for s in m["sections"]:
        assert s["residual"] < 0.02, s
        assert abs(s["R_eff"] - radius) < 1.0, s
        assert abs(s["u_c"] - u_max) < 0.1 * u_max, s
        assert abs(s["u_c_hagen_poiseuille"] - s["u_c"]) < 0.03 * s["u_c"], s
    assert m["G"] > 0
    assert m["flux_spread"] < 0.01, m["flux_spread"]
