"""GPU voxelizer and classification: bit-exact against the oracle and the
reference's PV2013 fixtures (ComplexGeometry.cpp:264-638)."""
import os

import numpy as np
import pytest

from test_oracle_voxelizer import tube_mesh

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def ctx():
    import pv2013_amd as P
    c = P.Context(0)
    yield c
    c.close()


def to_capi_planes(oracle_planes):
    import pv2013_amd as P
    out = []
    for p in oracle_planes:
        q = P.capi.PlaneDesc()
        for f, _ in P.capi.PlaneDesc._fields_:
            setattr(q, f, getattr(p, f))
        out.append(q)
    return out


def test_pv2013_bit_exact(ctx, oracle):
    import pv2013_amd as P
    z = np.load(os.path.join(ROOT, "tests", "golden", "pv2013.npz"))
    tri = z["tri"]
    g = P.grid_from_bbox(tri, 0.1)
    print("pv2013 voxelize", list(g.dims), flush=True)
    flags, ms = ctx.voxelize(tri, g)
    assert int(flags.sum()) == 1935409
    og = oracle.grid_from_bbox(oracle.bbox(tri), 0.1)
    assert np.array_equal(flags, oracle.voxelize(tri, og, 1))
    cl, _ = ctx.classify(g, flags, to_capi_planes(oracle.pv2013_planes()))
    for p, key in enumerate(["inlet", "out1", "out2", "out3", "out4", "out5", "out6"]):
        assert np.array_equal(np.nonzero(cl == 3 + p)[0], z[key]), key
    assert np.array_equal(np.nonzero(cl == 2)[0], z["wall"])
    assert np.array_equal(cl, oracle.classify(og, flags, oracle.pv2013_planes()))


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_random_tubes_match_bruteforce(ctx, oracle, seed):
    import pv2013_amd as P
    tri = tube_mesh(n_theta=10 + seed, n_z=5, jitter=0.2, seed=seed)
    if seed == 2:
        tri[:, :9] = np.round(tri[:, :9] * 4) / 4  # rays through shared edges and vertices
    if seed == 3:  # vertical facet with a tilted file normal: counts off its bbox
        tri = np.vstack([tri, [[1.0, 1.0, 1.0, 3.0, 1.0, 2.0, 5.0, 1.0, 3.0, 0.0, 0.9, 0.1]]])
    b = oracle.bbox(tri)
    for spacing in (0.5, 0.37):
        og = oracle.grid_from_bbox(b, spacing)
        g = P.grid_from_bbox(tri, spacing)
        assert list(g.dims) == list(og.dims)
        flags, _ = ctx.voxelize(tri, g)
        assert np.array_equal(flags, oracle.voxelize(tri, og, 0))


def facet_soup(oracle, seed):
    """A soup of facets whose file normals are not their geometric normals
    (random, tiny z component, exactly axial, powers of two), many of them
    placed so that an end of the t window falls on, or within rounding of, a
    grid point of some column.  The grid extends beyond the mesh, the spacing
    is 0.37 or 1 (integer vertices: rays through edges and vertices, offset
    columns on the ray box), and the window is [-1e-7, 1 + 1e-7] or [0, 1].
    Seeds 0 and 1 keep the moved vertices near the grid; seed 2 lets them go
    anywhere, so the ray box's z extent (the ray length) is about 1e12.
    Returns (tri, origin, spacing, dims, raybox, t_eps)."""
    rng = np.random.default_rng(100 + seed)
    n, dims = 400, (24, 20, 48)
    sp = 1.0 if seed == 1 else 0.37
    origin = np.array((-3.0, -3.0, -3.0) if seed == 1 else (-1.3, -0.7, -2.1))
    teps = 0.0 if seed == 1 else 1e-7
    ext = np.array(dims) * sp
    tri = np.zeros((n, 12))
    for v in range(3):
        tri[:, 3 * v:3 * v + 3] = origin + rng.uniform(-0.1, 1.1, (n, 3)) * ext
    if seed == 1:
        tri[:, :9] = np.round(tri[:, :9])
    kind = np.arange(n) % 5
    nrm = rng.uniform(-1, 1, (n, 3))
    nrm[kind == 1, 2] *= 1e-9
    nrm[kind == 2] = [0.0, 0.0, 1.0]
    nrm[kind == 3] = np.ldexp(1.0, rng.integers(-40, 3, (np.count_nonzero(kind == 3), 3)))
    tri[:, 9:] = nrm
    # the plane's height at one column on (or within 1e-13 of) a grid point
    for f in np.nonzero(rng.uniform(size=n) < 0.5)[0]:
        i, j, k = (int(rng.integers(0, d)) for d in dims)
        sx, sy = origin[0] + i * sp, origin[1] + j * sp
        nx, ny, nz = tri[f, 9:]
        a = (tri[f, 0] - sx) * nx + (tri[f, 1] - sy) * ny
        z = origin[2] + k * sp + rng.choice([0.0, 1e-13, -1e-13]) - a / nz
        if seed == 2 or abs(z - origin[2] - 0.5 * ext[2]) < ext[2]:
            tri[f, 2] = z
    return tri, tuple(origin), sp, dims, tuple(oracle.bbox(tri)), teps


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_facet_soup_with_file_normals(ctx, oracle, seed):
    """Flags of the facet soup equal the oracle's literal per-point count
    (ComplexGeometry.cpp:284-330) bit for bit."""
    import pv2013_amd as P
    tri, origin, sp, dims, rb, teps = facet_soup(oracle, seed)
    flags, _ = ctx.voxelize(tri, P.make_grid(origin, sp, dims, rb, t_eps=teps))
    ref = oracle.voxelize(tri, oracle.make_grid(origin, sp, dims, rb, t_eps=teps), 0)
    assert 0 < int(ref.sum()) < ref.size
    assert np.array_equal(flags, ref)


def test_c3_tube_blocks_match_oracle(ctx, oracle):
    """The seeded tortuous-tube mesh of the voxelizer benchmark, at 256^3 on the
    GPU: three sampled 16 x 16 column blocks equal the oracle's flags bit for
    bit (the oracle voxelizes each block as its own grid with the full ray box)."""
    from pv2013_amd import cases
    n = 256
    tri, g = cases.c3_mesh(seed=3, n=n)
    print("c3 tube mesh", len(tri), "facets, grid", list(g.dims), flush=True)
    assert tuple(g.origin) == (0.0, 0.0, 0.0) and g.spacing == 1.0
    flags, _ = ctx.voxelize(tri, g)
    vol = flags.reshape(n, n, n)
    # two blocks centred on columns that cross the tube, one at random
    cols = np.argwhere(vol.any(axis=0))  # (j, i) of columns holding fluid
    rng = np.random.default_rng(0)
    picks = [cols[k][::-1] - 8 for k in rng.choice(len(cols), 2, replace=False)]
    picks.append(rng.integers(0, n - 16, 2))
    fluid = 0
    for i0, j0 in (np.clip(p, 0, n - 16).astype(int) for p in picks):
        og = oracle.make_grid((float(i0), float(j0), 0.0), 1.0, (16, 16, n), tuple(g.raybox))
        sub = oracle.voxelize(tri, og, 2).reshape(n, 16, 16)
        mine = vol[:, j0:j0 + 16, i0:i0 + 16]
        assert np.array_equal(sub, mine), (i0, j0)
        fluid += int(sub.sum())
    assert fluid > 1000


def test_c3_tube_512_blocks_over_the_grid_match_oracle(ctx, oracle):
    """Config 3 at its stated size: the seeded tortuous tube (~544k facets) in
    the 512^3 grid.  Eight 16 x 16 column blocks spread over the whole 512^2
    grid equal the oracle's binned-mode flags bit for bit: the two corner
    blocks hold the columns on the ray box's faces (where the ray is shifted
    by 0.01, ComplexGeometry.cpp:334-357), three blocks sit on regular grid
    positions and three on columns that cross the tube."""
    from pv2013_amd import cases
    n = 512
    tri, g = cases.c3_mesh(seed=3, n=n)
    print("c3 tube mesh", len(tri), "facets, grid", list(g.dims), flush=True)
    flags, ms = ctx.voxelize(tri, g)
    print("voxelized in %.3f ms (kernel)" % ms, flush=True)
    vol = flags.reshape(n, n, n)
    cols = np.argwhere(vol.any(axis=0))  # (j, i) of columns holding fluid
    assert len(cols) > 1000
    order = np.argsort(cols[:, 0] * n + cols[:, 1])
    tube = [cols[order[int(f * (len(cols) - 1))]][::-1] - 8 for f in (0.2, 0.5, 0.8)]
    picks = [np.array([0, 0]), np.array([n - 16, n - 16]), np.array([128, 384]), np.array([256, 256]),
             np.array([384, 96])] + tube
    fluid = 0
    for i0, j0 in (np.clip(p, 0, n - 16).astype(int) for p in picks):
        print("block at column (%d, %d)" % (i0, j0), flush=True)
        og = oracle.make_grid((float(i0), float(j0), 0.0), 1.0, (16, 16, n), tuple(g.raybox))
        sub = oracle.voxelize(tri, og, 2).reshape(n, 16, 16)
        mine = vol[:, j0:j0 + 16, i0:i0 + 16]
        assert np.array_equal(sub, mine), (i0, j0)
        fluid += int(sub.sum())
    assert fluid > 10000
