"""Oracle voxelizer + classification against the reference's own fixtures and
against closed forms (ComplexGeometry.cpp:264-638, Subdomain.cpp:375-478)."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "pv2013.npz")


def test_pv2013_fixtures(oracle):
    z = np.load(GOLD)
    tri = z["tri"]
    assert tri.shape == (4988, 12)
    g = oracle.grid_from_bbox(oracle.bbox(tri), 0.1)
    assert list(g.dims) == list(z["dims"]) == [421, 351, 341]
    flags = oracle.voxelize(tri, g, mode=1)
    assert int(flags.sum()) == 1935409
    cl = oracle.classify(g, flags, oracle.pv2013_planes())
    for p, key in enumerate(["inlet", "out1", "out2", "out3", "out4", "out5", "out6"]):
        assert np.array_equal(np.nonzero(cl == 3 + p)[0], z[key]), key
    assert np.array_equal(np.nonzero(cl == 2)[0], z["wall"])


def tube_mesh(n_theta=12, n_z=6, r=3.3, z0=0.4, z1=8.6, cx=5.2, cy=4.9, caps=True, jitter=0.0, seed=0):
    """Closed polygonal tube along z with outward file normals."""
    rng = np.random.default_rng(seed)
    th = np.linspace(0, 2 * np.pi, n_theta, endpoint=False)
    zs = np.linspace(z0, z1, n_z)
    ring = lambda z: np.stack([cx + r * np.cos(th), cy + r * np.sin(th), np.full_like(th, z)], 1)
    rings = [ring(z) + jitter * rng.standard_normal((n_theta, 3)) * [1, 1, 0] for z in zs]
    tris = []
    for a in range(n_z - 1):
        for t in range(n_theta):
            u = (t + 1) % n_theta
            p0, p1, p2, p3 = rings[a][t], rings[a][u], rings[a + 1][u], rings[a + 1][t]
            tris += [(p0, p1, p2), (p0, p2, p3)]
    if caps:
        cb = np.array([cx, cy, z0]); ct = np.array([cx, cy, z1])
        for t in range(n_theta):
            u = (t + 1) % n_theta
            tris.append((cb, rings[0][u], rings[0][t]))
            tris.append((ct, rings[-1][t], rings[-1][u]))
    out = []
    for v in tris:
        v = np.array(v)
        n = np.cross(v[1] - v[0], v[2] - v[0])
        n = n / np.linalg.norm(n)
        out.append(np.concatenate([v.reshape(-1), n]))
    return np.array(out)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_modes_agree_with_literal_bruteforce(oracle, seed):
    tri = tube_mesh(jitter=0.15, seed=seed)
    # round vertices onto a coarse decimal grid so rays hit shared edges/vertices
    tri[:, :9] = np.round(tri[:, :9] * 4) / 4 if seed == 2 else tri[:, :9]
    g = oracle.grid_from_bbox(oracle.bbox(tri), 0.5)
    f0 = oracle.voxelize(tri, g, 0)
    f1 = oracle.voxelize(tri, g, 1)
    f2 = oracle.voxelize(tri, g, 2)
    assert f0.sum() > 0
    assert np.array_equal(f0, f1) and np.array_equal(f0, f2)


def test_degenerate_facets(oracle):
    """nz == 0 facets never count (t = +-inf or NaN); a vertical facet whose
    file normal has nz != 0 counts on its projected line, also outside its bbox."""
    base = tube_mesh(n_theta=8, n_z=3)
    vert = np.array([[1.0, 1.0, 1.0, 3.0, 1.0, 2.0, 5.0, 1.0, 3.0, 0.0, 1.0, 0.0]])
    tilt = np.array([[1.0, 1.0, 1.0, 3.0, 1.0, 2.0, 5.0, 1.0, 3.0, 0.0, 0.9, 0.1]])
    for extra in (vert, tilt):
        tri = np.vstack([base, extra])
        g = oracle.make_grid((0, 0, 0), 0.5, (21, 21, 21), oracle.bbox(tri))
        f0 = oracle.voxelize(tri, g, 0)
        assert np.array_equal(f0, oracle.voxelize(tri, g, 1))
        assert np.array_equal(f0, oracle.voxelize(tri, g, 2))
    gb = oracle.make_grid((0, 0, 0), 0.5, (21, 21, 21), oracle.bbox(np.vstack([base, vert])))
    assert np.array_equal(oracle.voxelize(base, gb, 0), oracle.voxelize(np.vstack([base, vert]), gb, 0))


def test_closed_cylinder_matches_analytic_inside(oracle):
    """Away from the surface the parity label equals the analytic inside test."""
    # centre off the grid lines so no column runs along a shared cap-fan edge,
    # which the inclusive edge test counts twice (the reference's behaviour)
    n = 64
    cx, cy = 8.1234, 7.9321
    tri = tube_mesh(n_theta=n, n_z=2, r=6.0, z0=1.0, z1=14.0, cx=cx, cy=cy)
    g = oracle.make_grid((0, 0, 0), 0.5, (33, 33, 33), (0.0, 16.0, 0.0, 16.0, 0.0, 16.0))
    fl = oracle.voxelize(tri, g, 1).reshape(33, 33, 33)  # [k][j][i]
    k, j, i = np.meshgrid(np.arange(33), np.arange(33), np.arange(33), indexing="ij")
    x, y, z = 0.5 * i, 0.5 * j, 0.5 * k
    r = np.hypot(x - cx, y - cy)
    r_in = 6.0 * np.cos(np.pi / n)  # inscribed radius of the polygon
    inside = (r < r_in - 0.05) & (z > 1.05) & (z < 13.95)
    outside = (r > 6.05) | (z < 0.95) | (z > 14.05)
    assert fl[inside].all()
    assert not fl[outside].any()


def test_classification_boundary_rule(oracle):
    """A filled box: boundary = fluid with an outside 26-neighbour or on the grid face."""
    dims = (7, 6, 5)
    flags = np.zeros(dims[::-1], np.uint8)
    flags[1:4, 1:5, 1:6] = 1
    g = oracle.make_grid((0, 0, 0), 1.0, dims, (0, 6, 0, 5, 0, 4))
    cl = oracle.classify(g, flags.reshape(-1), []).reshape(dims[::-1])
    assert (cl[2, 2:4, 2:5] == 1).all()
    assert cl[1:4, 1:5, 1:6].sum() - (cl[2, 2:4, 2:5] == 1).sum() == 2 * ((cl[1:4, 1:5, 1:6] == 2).sum())
    edge = flags.copy(); edge[:] = 1
    cl = oracle.classify(g, edge.reshape(-1), []).reshape(dims[::-1])
    assert (cl[1:-1, 1:-1, 1:-1] == 1).all() and (cl[0] == 2).all() and (cl[:, :, -1] == 2).all()
