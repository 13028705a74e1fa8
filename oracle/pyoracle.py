"""ctypes binding of the CPU oracle (oracle/liblbm_oracle.so).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() as the
checker and by bench.py's cpu_baseline leg.  Never imported by pv2013_amd.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liblbm_oracle.so")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "all"])


class Grid(C.Structure):
    _fields_ = [("origin", C.c_double * 3), ("spacing", C.c_double), ("dims", C.c_int32 * 3),
                ("pad_", C.c_int32), ("raybox", C.c_double * 6), ("ray_offset", C.c_double),
                ("ray_pad", C.c_double), ("t_eps", C.c_double)]


class Plane(C.Structure):
    _fields_ = [("axis", C.c_int32), ("nranges", C.c_int32), ("value_hi", C.c_double),
                ("value_lo", C.c_double), ("tol", C.c_double), ("range_axis", C.c_int32 * 3),
                ("pad_", C.c_int32), ("range_lo", C.c_double * 3), ("range_hi", C.c_double * 3)]


class BC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("axis", C.c_int32), ("sigma", C.c_int32),
                ("profile", C.c_int32), ("rho", C.c_double), ("u", C.c_double * 3),
                ("center", C.c_double * 2), ("radius", C.c_double), ("waveform", C.c_int32),
                ("typo", C.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        P = np.ctypeslib.ndpointer
        L.orc_lattice.argtypes = [C.c_int, P(np.int32), P(np.float64), P(np.int32)]
        L.orc_read_stl_ascii.argtypes = [C.c_char_p, C.c_void_p, C.c_int64]
        L.orc_read_stl_ascii.restype = C.c_int64
        L.orc_bbox.argtypes = [P(np.float64), C.c_int64, P(np.float64)]
        L.orc_grid_from_bbox.argtypes = [P(np.float64), C.c_double, C.POINTER(Grid)]
        L.orc_voxelize.argtypes = [P(np.float64), C.c_int64, C.POINTER(Grid), P(np.uint8), C.c_int]
        L.orc_split_long_double.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_classify.argtypes = [C.POINTER(Grid), P(np.uint8), C.c_void_p, C.c_int]
        L.orc_sim_create.argtypes = [C.c_int, P(np.int32), P(np.uint8), C.c_double, C.c_void_p,
                                     C.c_int, C.c_int]
        L.orc_sim_create.restype = C.c_void_p
        L.orc_sim_create_periodic.argtypes = [C.c_int, P(np.int32), P(np.uint8), C.c_double, C.c_void_p,
                                              C.c_int, C.c_int, C.c_int]
        L.orc_sim_create_periodic.restype = C.c_void_p
        L.orc_sim_destroy.argtypes = [C.c_void_p]
        L.orc_sim_set_waveform.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int,
                                           P(np.float64), P(np.float64), C.c_double]
        L.orc_sim_step.argtypes = [C.c_void_p, C.c_int64]
        L.orc_sim_step_passes.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_int32)]
        L.orc_sim_step_threads.argtypes = [C.c_void_p, C.c_int64, C.c_int]
        L.orc_sim_step_threads.restype = C.c_int
        L.orc_threads_available.restype = C.c_int
        L.orc_sim_nodes.argtypes = [C.c_void_p]
        L.orc_sim_nodes.restype = C.c_int64
        L.orc_sim_iter.argtypes = [C.c_void_p]
        L.orc_sim_iter.restype = C.c_int64
        for fn in ("orc_sim_get_f_dense", "orc_sim_set_f_dense", "orc_sim_get_wss_dense"):
            getattr(L, fn).argtypes = [C.c_void_p, P(np.float64)]
        L.orc_sim_get_macros_dense.argtypes = [C.c_void_p, P(np.float64), P(np.float64)]
        L.orc_sim_mass.argtypes = [C.c_void_p]
        L.orc_sim_mass.restype = C.c_double
        L.orc_waveform_eval.argtypes = [C.c_void_p, C.c_int, C.c_int64]
        L.orc_waveform_eval.restype = C.c_double
        L.orc_collide_node.argtypes = [C.c_int, C.c_double, P(np.float64), P(np.float64),
                                       P(np.float64), P(np.float64)]
        L.orc_zouhe_node.argtypes = [C.c_int, C.POINTER(BC), P(np.float64), P(np.float64)]
        _lib = L
    return _lib


def lattice(q):
    c = np.zeros(3 * q, np.int32)
    w = np.zeros(q, np.float64)
    opp = np.zeros(q, np.int32)
    assert lib().orc_lattice(q, c, w, opp) == 0
    return c.reshape(q, 3), w, opp


def read_stl_ascii(path):
    n = lib().orc_read_stl_ascii(path.encode(), None, 0)
    if n < 0:
        raise OSError(path)
    tri = np.zeros((n, 12), np.float64)
    lib().orc_read_stl_ascii(path.encode(), tri.ctypes.data, n)
    return tri


def bbox(tri):
    b = np.zeros(6)
    tri = np.ascontiguousarray(tri, np.float64)
    lib().orc_bbox(tri, len(tri), b)
    return b


def grid_from_bbox(b, spacing):
    g = Grid()
    lib().orc_grid_from_bbox(np.ascontiguousarray(b, np.float64), spacing, C.byref(g))
    return g


def make_grid(origin, spacing, dims, raybox, ray_offset=0.01, ray_pad=0.1, t_eps=1e-7):
    g = Grid()
    g.origin[:] = list(origin)
    g.spacing = spacing
    g.dims[:] = list(dims)
    g.raybox[:] = list(raybox)
    g.ray_offset, g.ray_pad, g.t_eps = ray_offset, ray_pad, t_eps
    return g


def voxelize(tri, grid, mode=1):
    tri = np.ascontiguousarray(tri, np.float64)
    n = grid.dims[0] * grid.dims[1] * grid.dims[2]
    flags = np.zeros(n, np.uint8)
    assert lib().orc_voxelize(tri, len(tri), C.byref(grid), flags, mode) == 0
    return flags


def split_long_double(text):
    hi, lo = C.c_double(), C.c_double()
    assert lib().orc_split_long_double(text.encode(), C.byref(hi), C.byref(lo)) == 0
    return hi.value, lo.value


def plane(axis, value_text, tol=1e-2, ranges=()):
    p = Plane()
    p.axis = axis
    p.value_hi, p.value_lo = split_long_double(value_text)
    p.tol = tol
    p.nranges = len(ranges)
    for r, (ax, lo, hi) in enumerate(ranges):
        p.range_axis[r] = ax
        p.range_lo[r] = lo
        p.range_hi[r] = hi
    return p


def classify(grid, flags, planes):
    arr = (Plane * max(1, len(planes)))(*planes)
    out = np.ascontiguousarray(flags, np.uint8).copy()
    assert lib().orc_classify(C.byref(grid), out, C.cast(arr, C.c_void_p), len(planes)) == 0
    return out


def make_bc(kind, axis, sigma, rho=1.0, u=(0.0, 0.0, 0.0), profile=0, center=(0.0, 0.0),
            radius=1.0, waveform=-1, typo=0):
    b = BC()
    b.kind, b.axis, b.sigma, b.profile = kind, axis, sigma, profile
    b.rho = rho
    b.u[:] = list(u)
    b.center[:] = list(center)
    b.radius = radius
    b.waveform = waveform
    b.typo = typo
    return b


class Sim:
    """Reference-semantics FP64 solver over a flag box (x-fastest sites)."""

    def __init__(self, q, dims, flags, tau, bcs=(), quirks=1, periodic=0):
        self.q = q
        self.dims = tuple(int(d) for d in dims)
        self.nsites = self.dims[0] * self.dims[1] * self.dims[2]
        arr = (BC * max(1, len(bcs)))(*bcs)
        self._bcs = arr
        self.h = lib().orc_sim_create_periodic(q, np.array(self.dims, np.int32),
                                               np.ascontiguousarray(flags, np.uint8), tau,
                                               C.cast(arr, C.c_void_p), len(bcs), quirks, periodic)
        if not self.h:
            raise ValueError("orc_sim_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sim_destroy(self.h)
            self.h = None

    @property
    def nodes(self):
        return lib().orc_sim_nodes(self.h)

    @property
    def iter(self):
        return lib().orc_sim_iter(self.h)

    def set_waveform(self, slot, kind, period, a=(), phi=(), dt=1.0):
        a = np.ascontiguousarray(a if len(a) else [0.0], np.float64)
        phi = np.ascontiguousarray(phi if len(phi) else [0.0], np.float64)
        assert lib().orc_sim_set_waveform(self.h, slot, kind, period, len(a) if kind == 1 else 0,
                                          a, phi, dt) == 0

    def step(self, n=1):
        lib().orc_sim_step(self.h, n)

    def step_passes(self, n, order):
        """n steps with the boundary lists as separate passes in `order`
        (0 inlet, 1 wall, 2 outlet), as the reference runs them."""
        o = (C.c_int32 * 3)(*order)
        assert lib().orc_sim_step_passes(self.h, n, o) == 0

    def step_threads(self, n=1, threads=0):
        """The same steps on OpenMP threads (0: the OpenMP default)."""
        assert lib().orc_sim_step_threads(self.h, n, threads) == 0

    def f(self):
        out = np.zeros(self.q * self.nsites)
        lib().orc_sim_get_f_dense(self.h, out)
        return out.reshape(self.q, self.nsites)

    def set_f(self, f):
        lib().orc_sim_set_f_dense(self.h, np.ascontiguousarray(f, np.float64).reshape(-1))

    def macros(self):
        rho = np.zeros(self.nsites)
        u = np.zeros(3 * self.nsites)
        lib().orc_sim_get_macros_dense(self.h, rho, u)
        return rho, u.reshape(3, self.nsites)

    def wss(self):
        out = np.zeros(self.nsites)
        lib().orc_sim_get_wss_dense(self.h, out)
        return out

    def mass(self):
        return lib().orc_sim_mass(self.h)


def threads_available():
    return lib().orc_threads_available()


def collide_node(q, omega, f):
    f = np.ascontiguousarray(f, np.float64)
    fnew = np.zeros(q)
    rho = np.zeros(1)
    u = np.zeros(3)
    lib().orc_collide_node(q, omega, f, fnew, rho, u)
    return fnew, rho[0], u


def zouhe_node(q, bc, uvec, f):
    f = np.ascontiguousarray(f, np.float64).copy()
    lib().orc_zouhe_node(q, C.byref(bc), np.ascontiguousarray(uvec, np.float64), f)
    return f


# PV2013 inlet/outlet planes, ComplexGeometry.cpp:579-633 (axis 0=x, 1=y, 2=z)
def pv2013_planes():
    return [
        plane(1, "-175.275", ranges=((2, -129.979, -124.799), (0, -9.9, -3.3))),
        plane(1, "-164.075"),
        plane(1, "-174.475", ranges=((0, -5.4, -5.0), (2, -101.379, -100.979))),
        plane(0, "-28.3", ranges=((2, -97.2794, -96.2794), (1, -179.975, -178.875))),
        plane(0, "-25.5", ranges=((1, -180.675, -179.875), (2, -100.779, -100.379))),
        plane(0, "-37.0"),
        plane(2, "-108.379", ranges=((1, -196.375, -194.575), (0, -29.7, -28.1))),
    ]
