"""ctypes binding of libpv2013_lbm.so (include/lbm_capi.h).

The library is the product: there is no Python or CPU fallback.  Loading fails
loudly when the shared library has not been built.
"""
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "csrc", "libpv2013_lbm.so")

LBM_OK = 0
STORE_F64, STORE_F32, STORE_F16S = 0, 1, 2
SCHEME_TWO_LATTICE, SCHEME_AA = 0, 1
OPT_NO_ROW_FASTPATH = 1
OPT_NO_GRAPH = 2
OPT_PERIODIC_X, OPT_PERIODIC_Y, OPT_PERIODIC_Z = 4, 8, 16
OPT_GENERIC_KERNEL = 32  # one thread per node through the site map
OPT_SPLIT_LAUNCH = 64  # records and row segments in two launches
OPT_WIDE_RECORDS = 128  # FP16 chunks with records through 64-bit addresses (automatic >= 4 GiB)
PROBE_TWO, PROBE_IN_PLACE, PROBE_TWO_SHIFTED, PROBE_IN_PLACE_SHIFTED, PROBE_AA = 0, 1, 2, 3, 4  # lbm_ctx_bandwidth_lattice
OPT_ONE_CHUNK_PER_WAVE = 256  # two-lattice row chunks one per wavefront instead of two-chunk windows
OPT_FORWARD_ONLY = 512  # every step launch walks the lattice upward (default: launches alternate direction)
BC_VELOCITY, BC_PRESSURE = 0, 1
PROFILE_UNIFORM, PROFILE_PARABOLIC = 0, 1
WAVE_NONE, WAVE_FOURIER, WAVE_PV2013 = 0, 1, 2
STORAGE_BYTES = {STORE_F64: 8, STORE_F32: 4, STORE_F16S: 2}
STORAGE_NAMES = {STORE_F64: "f64", STORE_F32: "f32", STORE_F16S: "f16s"}

EXPORTS = (
    "lbm_ctx_create", "lbm_ctx_destroy", "lbm_last_error", "lbm_version", "lbm_read_stl_ascii",
    "lbm_read_stl_binary", "lbm_grid_from_bbox", "lbm_voxelize", "lbm_classify", "lbm_domain_create",
    "lbm_domain_destroy", "lbm_domain_info", "lbm_set_waveform", "lbm_step", "lbm_step_timed",
    "lbm_sync", "lbm_iteration", "lbm_get_f", "lbm_set_f", "lbm_get_macros", "lbm_get_wss",
    "lbm_reduce", "lbm_get_macros_box", "lbm_get_wss_box", "lbm_batch_create", "lbm_autotune",
    "lbm_domain_get_stats", "lbm_ctx_bandwidth", "lbm_ctx_bandwidth_lattice", "lbm_voxelize_dev",
)


class LbmError(RuntimeError):
    pass


class GridDesc(C.Structure):
    _fields_ = [("origin", C.c_double * 3), ("spacing", C.c_double), ("dims", C.c_int32 * 3),
                ("pad_", C.c_int32), ("raybox", C.c_double * 6), ("ray_offset", C.c_double),
                ("ray_pad", C.c_double), ("t_eps", C.c_double)]


class PlaneDesc(C.Structure):
    _fields_ = [("axis", C.c_int32), ("nranges", C.c_int32), ("value_hi", C.c_double),
                ("value_lo", C.c_double), ("tol", C.c_double), ("range_axis", C.c_int32 * 3),
                ("pad_", C.c_int32), ("range_lo", C.c_double * 3), ("range_hi", C.c_double * 3)]


class BcDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("axis", C.c_int32), ("sigma", C.c_int32),
                ("profile", C.c_int32), ("rho", C.c_double), ("u", C.c_double * 3),
                ("center", C.c_double * 2), ("radius", C.c_double), ("waveform", C.c_int32),
                ("typo", C.c_int32)]


class DomainDesc(C.Structure):
    _fields_ = [("dims", C.c_int32 * 3), ("q", C.c_int32), ("storage", C.c_int32),
                ("scheme", C.c_int32), ("quirks", C.c_int32), ("block", C.c_int32),
                ("tau", C.c_double), ("flags", C.c_void_p), ("bcs", C.c_void_p),
                ("nbc", C.c_int32), ("fast_axis", C.c_int32), ("options", C.c_int32)]


class DomainStats(C.Structure):
    _fields_ = [("nodes", C.c_int64), ("boundary_records", C.c_int64), ("wall_nodes", C.c_int64),
                ("bc_nodes", C.c_int64), ("waves", C.c_int64), ("regular_waves", C.c_int64),
                ("segments", C.c_int64), ("device_bytes", C.c_int64), ("fast_axis", C.c_int32),
                ("scheme", C.c_int32), ("layout", C.c_int32), ("block", C.c_int32), ("step_waves", C.c_int64)]


class TuneResult(C.Structure):
    _fields_ = [("scheme", C.c_int32), ("fast_axis", C.c_int32), ("options", C.c_int32),
                ("ntried", C.c_int32), ("mlups", C.c_double), ("key", C.c_uint64), ("block", C.c_int32),
                ("pad_", C.c_int32)]


class BatchCase(C.Structure):
    _fields_ = [("origin", C.c_int32 * 3), ("dims", C.c_int32 * 3), ("bc_base", C.c_int32),
                ("wave_base", C.c_int32)]


_lib = None


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LbmError("libpv2013_lbm.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, i64, i32, dbl = C.c_void_p, C.c_int64, C.c_int, C.c_double
        P = C.POINTER
        L.lbm_ctx_create.argtypes = [i32, P(vp)]
        L.lbm_ctx_destroy.argtypes = [vp]
        L.lbm_ctx_destroy.restype = None
        L.lbm_last_error.argtypes = [vp]
        L.lbm_last_error.restype = C.c_char_p
        L.lbm_read_stl_ascii.argtypes = [C.c_char_p, vp, i64]
        L.lbm_read_stl_ascii.restype = i64
        L.lbm_read_stl_binary.argtypes = [C.c_char_p, vp, i64]
        L.lbm_read_stl_binary.restype = i64
        L.lbm_grid_from_bbox.argtypes = [vp, i64, dbl, P(GridDesc)]
        L.lbm_voxelize.argtypes = [vp, vp, i64, P(GridDesc), vp, P(dbl)]
        L.lbm_voxelize_dev.argtypes = [vp, vp, i64, P(GridDesc), vp, P(dbl)]
        L.lbm_classify.argtypes = [vp, P(GridDesc), vp, i32, vp, P(dbl)]
        L.lbm_domain_create.argtypes = [vp, P(DomainDesc), P(vp)]
        L.lbm_domain_destroy.argtypes = [vp]
        L.lbm_domain_destroy.restype = None
        L.lbm_domain_info.argtypes = [vp, P(i64), P(i64), P(i64)]
        L.lbm_set_waveform.argtypes = [vp, i32, i32, dbl, i32, vp, vp, dbl]
        L.lbm_step.argtypes = [vp, i64]
        L.lbm_step_timed.argtypes = [vp, i64, P(dbl)]
        L.lbm_sync.argtypes = [vp]
        L.lbm_iteration.argtypes = [vp]
        L.lbm_iteration.restype = i64
        for fn in ("lbm_get_f", "lbm_set_f", "lbm_get_wss"):
            getattr(L, fn).argtypes = [vp, vp]
        L.lbm_get_macros.argtypes = [vp, vp, vp]
        L.lbm_reduce.argtypes = [vp, P(dbl), vp]
        L.lbm_get_macros_box.argtypes = [vp, vp, vp, vp, vp]
        L.lbm_get_wss_box.argtypes = [vp, vp, vp, vp]
        L.lbm_batch_create.argtypes = [vp, vp, i32, i32, P(vp), vp]
        L.lbm_domain_get_stats.argtypes = [vp, P(DomainStats)]
        L.lbm_autotune.argtypes = [vp, P(DomainDesc), i32, C.c_char_p, P(TuneResult)]
        L.lbm_ctx_bandwidth.argtypes = [vp, i64, i32, i32, P(dbl)]
        L.lbm_ctx_bandwidth_lattice.argtypes = [vp, i32, i32, i64, i32, i32, P(dbl)]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def read_stl(path):
    """Read an ASCII or binary STL (GeometryReader.cpp:33-176) into (n, 12) doubles."""
    L = lib()
    with open(path, "rb") as fh:
        head = fh.read(80)
    binary = not head.lstrip().lower().startswith(b"solid")
    fn = L.lbm_read_stl_binary if binary else L.lbm_read_stl_ascii
    n = fn(path.encode(), None, 0)
    if n <= 0 and not binary:
        fn = L.lbm_read_stl_binary
        n = fn(path.encode(), None, 0)
    if n < 0:
        raise LbmError("cannot read %s (%d)" % (path, n))
    tri = np.zeros((n, 12), np.float64)
    fn(path.encode(), _ptr(tri), n)
    return tri


def grid_from_bbox(tri, spacing):
    tri = np.ascontiguousarray(tri, np.float64)
    g = GridDesc()
    rc = lib().lbm_grid_from_bbox(_ptr(tri), len(tri), spacing, C.byref(g))
    if rc:
        raise LbmError("lbm_grid_from_bbox failed (%d)" % rc)
    return g


def make_grid(origin, spacing, dims, raybox, ray_offset=0.01, ray_pad=0.1, t_eps=1e-7):
    g = GridDesc()
    g.origin[:] = list(origin)
    g.spacing = spacing
    g.dims[:] = [int(v) for v in dims]
    g.raybox[:] = list(raybox)
    g.ray_offset, g.ray_pad, g.t_eps = ray_offset, ray_pad, t_eps
    return g


def split_long_double(text):
    """Long double value of a decimal literal as hi + lo doubles."""
    v = np.longdouble(text)
    hi = float(v)
    lo = float(v - np.longdouble(hi))
    return hi, lo


def make_plane(axis, value_text, tol=1e-2, ranges=()):
    p = PlaneDesc()
    p.axis = axis
    p.value_hi, p.value_lo = split_long_double(value_text)
    p.tol = tol
    p.nranges = len(ranges)
    for r, (ax, lo, hi) in enumerate(ranges):
        p.range_axis[r], p.range_lo[r], p.range_hi[r] = ax, lo, hi
    return p


def make_bc(kind, axis, sigma, rho=1.0, u=(0.0, 0.0, 0.0), profile=PROFILE_UNIFORM,
            center=(0.0, 0.0), radius=1.0, waveform=-1, typo=0):
    b = BcDesc()
    b.kind, b.axis, b.sigma, b.profile = kind, axis, sigma, profile
    b.rho = rho
    b.u[:] = list(u)
    b.center[:] = list(center)
    b.radius = radius
    b.waveform = waveform
    b.typo = typo
    return b


class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        rc = lib().lbm_ctx_create(device, C.byref(h))
        if rc:
            raise LbmError("lbm_ctx_create(%d) failed (%d)" % (device, rc))
        self.h = h
        self.domains = weakref.WeakSet()

    def close(self):
        """Destroys the context and every domain still open on it."""
        if getattr(self, "h", None):
            for d in list(self.domains):
                d.h = None  # freed by lbm_ctx_destroy below
            lib().lbm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def check(self, rc, what):
        if rc:
            msg = lib().lbm_last_error(self.h).decode(errors="replace")
            raise LbmError("%s failed (%d): %s" % (what, rc, msg))

    def bandwidth(self, nbytes, in_place=True, reps=20):
        """Measured streaming bandwidth in GB/s (bytes read + written per
        second) over a scratch buffer of nbytes (lbm_ctx_bandwidth)."""
        g = C.c_double()
        self.check(lib().lbm_ctx_bandwidth(self.h, int(nbytes) // 16 * 16, 1 if in_place else 0, reps, C.byref(g)),
                   "lbm_ctx_bandwidth")
        return g.value

    def bandwidth_lattice(self, q, elem_bytes, nodes, pattern=1, reps=20):
        """Measured GB/s of a step's traffic with no arithmetic: q arrays of
        nodes elements, read then written (lbm_ctx_bandwidth_lattice; pattern
        PROBE_TWO, PROBE_IN_PLACE, PROBE_TWO_SHIFTED, PROBE_IN_PLACE_SHIFTED or
        PROBE_AA)."""
        g = C.c_double()
        self.check(lib().lbm_ctx_bandwidth_lattice(self.h, q, elem_bytes, int(nodes), int(pattern), reps,
                                                   C.byref(g)), "lbm_ctx_bandwidth_lattice")
        return g.value

    def voxelize(self, tri, grid):
        """Geometry::labelPoints on the GPU: uint8 0/1 flags, x fastest.
        Returns (flags, kernel_ms)."""
        tri = np.ascontiguousarray(tri, np.float64)
        n = grid.dims[0] * grid.dims[1] * grid.dims[2]
        flags = np.zeros(n, np.uint8)
        ms = C.c_double()
        self.check(lib().lbm_voxelize(self.h, _ptr(tri), len(tri), C.byref(grid), _ptr(flags),
                                      C.byref(ms)), "lbm_voxelize")
        return flags, ms.value

    def classify(self, grid, flags, planes):
        arr = (PlaneDesc * max(1, len(planes)))(*planes)
        out = np.ascontiguousarray(flags, np.uint8).copy()
        ms = C.c_double()
        self.check(lib().lbm_classify(self.h, C.byref(grid), C.cast(arr, C.c_void_p), len(planes),
                                      _ptr(out), C.byref(ms)), "lbm_classify")
        return out, ms.value


def domain_desc(dims, flags, tau, bcs=(), q=19, storage=STORE_F64, scheme=SCHEME_TWO_LATTICE,
                quirks=1, block=0, fast_axis=-1, options=0):
    """A DomainDesc plus the Python objects that must outlive it."""
    dims = tuple(int(v) for v in dims)
    flags = np.ascontiguousarray(flags, np.uint8).reshape(-1)
    assert flags.size == dims[0] * dims[1] * dims[2]
    arr = (BcDesc * max(1, len(bcs)))(*bcs)
    d = DomainDesc()
    d.dims[:] = list(dims)
    d.q, d.storage, d.scheme, d.quirks, d.block = q, storage, scheme, quirks, block
    d.tau = tau
    d.flags = _ptr(flags)
    d.bcs = C.cast(arr, C.c_void_p)
    d.nbc = len(bcs)
    d.fast_axis = fast_axis
    d.options = options
    return d, (flags, arr)


class Domain:
    """One fluid domain on the GPU (SubDomain + Simulation of the reference)."""

    def __init__(self, ctx, dims, flags, tau, bcs=(), q=19, storage=STORE_F64,
                 scheme=SCHEME_TWO_LATTICE, quirks=1, block=0, fast_axis=-1, options=0, _handle=None):
        self.ctx = ctx
        self.q = q
        self.storage = storage
        self.dims = tuple(int(d) for d in dims)
        self.nsites = self.dims[0] * self.dims[1] * self.dims[2]
        self.nbc = len(bcs)
        ctx.domains.add(self)
        if _handle is not None:
            self.h = _handle
            return
        d, self._keep = domain_desc(dims, flags, tau, bcs, q, storage, scheme, quirks, block,
                                    fast_axis, options)
        h = C.c_void_p()
        ctx.check(lib().lbm_domain_create(ctx.h, C.byref(d), C.byref(h)), "lbm_domain_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            if getattr(self.ctx, "h", None):
                lib().lbm_domain_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def info(self):
        n, nb, by = C.c_int64(), C.c_int64(), C.c_int64()
        lib().lbm_domain_info(self.h, C.byref(n), C.byref(nb), C.byref(by))
        return {"nodes": n.value, "boundary_records": nb.value, "device_bytes": by.value}

    def stats(self):
        st = DomainStats()
        self.ctx.check(lib().lbm_domain_get_stats(self.h, C.byref(st)), "lbm_domain_get_stats")
        return {f: getattr(st, f) for f, _ in DomainStats._fields_}

    def set_waveform(self, slot, kind, period, a=(), phi=(), dt=1.0):
        a = np.ascontiguousarray(a if len(a) else [0.0], np.float64)
        phi = np.ascontiguousarray(phi if len(phi) else [0.0], np.float64)
        nk = len(a) if kind == WAVE_FOURIER else 0
        self.ctx.check(lib().lbm_set_waveform(self.h, slot, kind, period, nk, _ptr(a), _ptr(phi), dt),
                       "lbm_set_waveform")

    def step(self, n=1):
        self.ctx.check(lib().lbm_step(self.h, n), "lbm_step")

    def step_timed(self, n):
        ms = C.c_double()
        self.ctx.check(lib().lbm_step_timed(self.h, n, C.byref(ms)), "lbm_step_timed")
        return ms.value

    def sync(self):
        self.ctx.check(lib().lbm_sync(self.h), "lbm_sync")

    @property
    def iteration(self):
        return lib().lbm_iteration(self.h)

    def f(self):
        out = np.zeros(self.q * self.nsites)
        self.ctx.check(lib().lbm_get_f(self.h, _ptr(out)), "lbm_get_f")
        return out.reshape(self.q, self.nsites)

    def set_f(self, f):
        f = np.ascontiguousarray(f, np.float64).reshape(-1)
        self.ctx.check(lib().lbm_set_f(self.h, _ptr(f)), "lbm_set_f")

    def macros(self):
        rho = np.zeros(self.nsites)
        u = np.zeros(3 * self.nsites)
        self.ctx.check(lib().lbm_get_macros(self.h, _ptr(rho), _ptr(u)), "lbm_get_macros")
        return rho, u.reshape(3, self.nsites)

    def wss(self):
        out = np.zeros(self.nsites)
        self.ctx.check(lib().lbm_get_wss(self.h, _ptr(out)), "lbm_get_wss")
        return out

    def macros_box(self, origin, dims):
        o = np.ascontiguousarray(origin, np.int32)
        bd = np.ascontiguousarray(dims, np.int32)
        nb = int(np.prod(bd))
        rho = np.zeros(nb)
        u = np.zeros(3 * nb)
        self.ctx.check(lib().lbm_get_macros_box(self.h, _ptr(o), _ptr(bd), _ptr(rho), _ptr(u)),
                       "lbm_get_macros_box")
        return rho, u.reshape(3, nb)

    def wss_box(self, origin, dims):
        o = np.ascontiguousarray(origin, np.int32)
        bd = np.ascontiguousarray(dims, np.int32)
        out = np.zeros(int(np.prod(bd)))
        self.ctx.check(lib().lbm_get_wss_box(self.h, _ptr(o), _ptr(bd), _ptr(out)), "lbm_get_wss_box")
        return out

    def reduce(self):
        mass = C.c_double()
        flux = np.zeros(max(1, self.nbc))
        self.ctx.check(lib().lbm_reduce(self.h, C.byref(mass), _ptr(flux)), "lbm_reduce")
        return mass.value, flux[: self.nbc]


def batch(ctx, cases, stack_axis=1):
    """Pack several cases (each a dict of Domain arguments: dims, flags, tau,
    bcs, q, storage, quirks, ...) into one domain.  Returns (Domain, info)
    where info[c] = {"origin", "dims", "bc_base", "wave_base"}."""
    descs, keep = [], []
    for c in cases:
        d, k = domain_desc(**c)
        descs.append(d)
        keep.append(k)
    arr = (DomainDesc * len(descs))(*descs)
    info = (BatchCase * len(descs))()
    h = C.c_void_p()
    ctx.check(lib().lbm_batch_create(ctx.h, C.cast(arr, C.c_void_p), len(descs), stack_axis, C.byref(h),
                                     C.cast(info, C.c_void_p)), "lbm_batch_create")
    out = [{"origin": tuple(i.origin), "dims": tuple(i.dims), "bc_base": i.bc_base,
            "wave_base": i.wave_base} for i in info]
    dims = [0, 0, 0]
    for i in out:
        for a in range(3):
            dims[a] = max(dims[a], i["origin"][a] + i["dims"][a])
    nbc = sum(len(c.get("bcs", ())) for c in cases)
    dom = Domain(ctx, dims, None, cases[0]["tau"], [None] * nbc, q=cases[0].get("q", 19),
                 storage=cases[0].get("storage", STORE_F64), _handle=h)
    return dom, out


def autotune(ctx, steps=200, cache_path=None, **domain_kw):
    """Fastest layout for a case (lbm_autotune).  Returns a dict with scheme,
    fast_axis, options, block, mlups, ntried and the cache key."""
    d, keep = domain_desc(**domain_kw)
    r = TuneResult()
    ctx.check(lib().lbm_autotune(ctx.h, C.byref(d), steps,
                                 cache_path.encode() if cache_path else None, C.byref(r)), "lbm_autotune")
    return {"scheme": r.scheme, "fast_axis": r.fast_axis, "options": r.options, "block": r.block,
            "mlups": r.mlups, "ntried": r.ntried, "key": "%016x" % r.key}
