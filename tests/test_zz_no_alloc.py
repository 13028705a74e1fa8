"""lbm_voxelize_dev, the voxelizer entry point that allocates nothing (caller
device buffers in and out, its bins in the library's fixed 64 MiB scratch
array), against lbm_voxelize (the allocating entry point) bit for bit: two
calls in a row, the second on a smaller grid and mesh, so the second call
starts from the first call's scratch contents; and an oversized grid is
rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def hip_runtime(P):
    """The HIP runtime instance libpv2013_lbm.so is linked against (the one
    already mapped into this process), so buffers allocated here are the
    library's kind of device memory."""
    P.lib()
    with open("/proc/self/maps") as fh:
        paths = {line.split()[-1] for line in fh if "libamdhip64" in line and line.split()[-1].startswith("/")}
    ours = [p for p in paths if "/torch/" not in p]
    assert ours, paths
    hip = C.CDLL(ours[0])
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(1, nbytes)) == 0

    def close(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = C.c_void_p()


def voxelize_dev(P, hip, ctx, tri, grid):
    L = P.lib()
    tri = np.ascontiguousarray(tri, np.float64)
    n = grid.dims[0] * grid.dims[1] * grid.dims[2]
    dt, df = DevBuf(hip, tri.nbytes), DevBuf(hip, n)
    try:
        assert hip.hipMemcpy(dt.p, tri.ctypes.data_as(C.c_void_p), tri.nbytes, H2D) == 0
        ms = C.c_double()
        ctx.check(L.lbm_voxelize_dev(ctx.h, dt.p, len(tri), C.byref(grid), df.p, C.byref(ms)), "lbm_voxelize_dev")
        out = np.zeros(n, np.uint8)
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), df.p, n, D2H) == 0
        return out, ms.value
    finally:
        dt.close()
        df.close()


def test_voxelize_dev_matches_voxelize_on_two_calls(oracle):
    import pv2013_amd as P
    from pv2013_amd import cases
    from test_oracle_voxelizer import tube_mesh
    ctx = P.Context(0)
    hip = hip_runtime(P)
    big_tri, big_g = cases.c3_mesh(seed=3, n=160, h=2.5)
    small_tri = tube_mesh(n_theta=12, n_z=5, jitter=0.2, seed=1)
    small_g = P.grid_from_bbox(small_tri, 0.37)
    for name, tri, g in (("tube 160^3", big_tri, big_g), ("small tube", small_tri, small_g)):
        print("lbm_voxelize_dev:", name, len(tri), "facets, grid", list(g.dims), flush=True)
        want, _ = ctx.voxelize(tri, g)
        got, ms = voxelize_dev(P, hip, ctx, tri, g)
        assert 0 < int(want.sum()) < want.size
        assert np.array_equal(got, want), name
        assert ms > 0
    # 20000 x 20000 columns are 2500^2 tiles, whose three int32 tile arrays
    # alone take 75 MB: above the 64 MiB scratch, rejected before any launch
    huge = P.make_grid((0.0, 0.0, 0.0), 1.0, (20000, 20000, 1), (0.0, 19999.0, 0.0, 19999.0, 0.0, 0.0))
    with pytest.raises(P.capi.LbmError) as e:
        voxelize_dev(P, hip, ctx, small_tri, huge)
    assert "(-34)" in str(e.value) and "scratch" in str(e.value), str(e.value)
    # the context still works
    got, _ = voxelize_dev(P, hip, ctx, small_tri, small_g)
    assert np.array_equal(got, ctx.voxelize(small_tri, small_g)[0])
    ctx.close()
