"""GPU stream-collide engine against the CPU oracle.

FP64: bit-identical to the oracle (the reference's arithmetic) wherever no
transcendental function is involved; with a Fourier inlet waveform (sin on the
GPU vs glibc) within 1e-12 relative.  FP32 / FP16 shifted storage: velocity
error against the FP64 GPU run on the same case, with the stated bounds."""
import numpy as np
import pytest

from lbm_cases import box_flags, pipe_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def P():
    import pv2013_amd
    return pv2013_amd


@pytest.fixture(scope="module")
def ctx(P):
    c = P.Context(0)
    yield c
    c.close()


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


def both(P, oracle, ctx, q, dims, flags, tau, bcs, quirks=1, storage=0, fast_axis=-1):
    dom = P.Domain(ctx, dims, flags, tau, [P.make_bc(**b) for b in bcs], q=q, storage=storage,
                   quirks=quirks, fast_axis=fast_axis)
    ref = oracle.Sim(q, dims, flags, tau, [oracle.make_bc(**b) for b in bcs], quirks=quirks)
    return dom, ref


def six_face_box(oracle, dims):
    """Fluid box touching all six grid faces: every face orientation carries a
    Zou-He condition (velocity and pressure, with the x-face typo)."""
    flags = np.ones(dims[0] * dims[1] * dims[2], np.uint8)
    n = dims
    planes = [oracle.plane(1, "0"), oracle.plane(2, str(n[2] - 1)), oracle.plane(0, "0"),
              oracle.plane(0, str(n[0] - 1)), oracle.plane(1, str(n[1] - 1)), oracle.plane(2, "0")]
    g = oracle.make_grid((0, 0, 0), 1.0, dims, (0, n[0] - 1, 0, n[1] - 1, 0, n[2] - 1))
    cl = oracle.classify(g, flags, planes)
    bcs = [dict(kind=0, axis=1, sigma=1, u=(0.004, 0.02, -0.003)),
           dict(kind=0, axis=2, sigma=-1, u=(0.0, 0.0, -0.015), profile=1,
                center=((n[0] - 1) / 2, (n[1] - 1) / 2), radius=n[0] / 2.0),
           dict(kind=1, axis=0, sigma=1, rho=1.01, typo=1),
           dict(kind=1, axis=0, sigma=-1, rho=1.0),
           dict(kind=1, axis=1, sigma=-1, rho=0.995),
           dict(kind=1, axis=2, sigma=1, rho=1.005)]
    return cl, bcs


def near_eq_state(oracle, q, flags, rng, amp=0.01):
    c, w, _ = oracle.lattice(q)
    ns = flags.size
    rho = 1 + amp * rng.standard_normal(ns)
    u = 0.02 * rng.standard_normal((3, ns))
    cu = c @ u
    usq = (u * u).sum(0)
    f = w[:, None] * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq)
    f *= 1 + amp * rng.standard_normal(f.shape)
    f[:, flags == 0] = 0.0
    return f


@pytest.mark.parametrize("q", [19, 27])
@pytest.mark.parametrize("quirks", [1, 0])
def test_fp64_bitwise_all_faces(P, oracle, ctx, q, quirks):
    dims = (10, 11, 12)
    cl, bcs = six_face_box(oracle, dims)
    print("six-face box q=%d quirks=%d" % (q, quirks), flush=True)
    dom, ref = both(P, oracle, ctx, q, dims, cl, 0.7, bcs, quirks=quirks)
    f0 = near_eq_state(oracle, q, cl, np.random.default_rng(q + quirks))
    dom.set_f(f0)
    ref.set_f(f0)
    assert bits_equal(dom.f(), ref.f())
    for n in (1, 9, 30):
        dom.step(n)
        ref.step(n)
        assert bits_equal(dom.f(), ref.f()), "step %d" % ref.iter
    r1, u1 = dom.macros()
    r2, u2 = ref.macros()
    assert bits_equal(r1, r2) and bits_equal(u1, u2)
    assert bits_equal(dom.wss(), ref.wss())


@pytest.mark.parametrize("fast_axis", [-1, 0, 1])
def test_fp64_bitwise_pipe_config1(P, oracle, ctx, fast_axis):
    """D3Q19 FP64, 32x32x128 z-pipe of radius 12, tau 0.8, parabolic inlet
    u_max 0.05, outlet rho 1.0, 2000 steps from rest (the small plain case)."""
    dims = (32, 32, 128)
    cl, bcs = pipe_case(oracle, dims, 12.0, 0.05)
    steps = 2000 if fast_axis == -1 else 200
    print("pipe 32x32x128 fast_axis=%d %d steps" % (fast_axis, steps), flush=True)
    dom, ref = both(P, oracle, ctx, 19, dims, cl, 0.8, bcs, fast_axis=fast_axis)
    dom.step(steps)
    ref.step(steps)
    assert bits_equal(dom.f(), ref.f())
    assert bits_equal(dom.wss(), ref.wss())
    m, flux = dom.reduce()
    rho, _ = ref.macros()
    assert abs(m - rho.sum()) < 1e-9 * m


def test_fp64_waveform_within_1e12(P, oracle, ctx):
    dims = (16, 16, 40)
    cl, bcs = pipe_case(oracle, dims, 6.0, 0.03, waveform=0)
    dom, ref = both(P, oracle, ctx, 19, dims, cl, 0.8, bcs)
    rng = np.random.default_rng(4)
    a, phi = 0.3 * rng.random(4), 2 * np.pi * rng.random(4)
    dom.set_waveform(0, P.WAVE_FOURIER, 400.0, a, phi)
    ref.set_waveform(0, 1, 400.0, a, phi)
    dom.step(300)
    ref.step(300)
    r1, u1 = dom.macros()
    r2, u2 = ref.macros()
    assert np.abs(r1 - r2).max() <= 1e-12 * np.abs(r2).max()
    assert np.linalg.norm(u1 - u2) <= 1e-12 * np.linalg.norm(u2)


@pytest.mark.parametrize("storage,bound", [(1, 1e-5), (2, 5e-3)])
def test_low_precision_error_vs_fp64(P, oracle, ctx, storage, bound):
    dims = (32, 32, 128)
    cl, bcs = pipe_case(oracle, dims, 12.0, 0.05)
    bc = [P.make_bc(**b) for b in bcs]
    ref = P.Domain(ctx, dims, cl, 0.8, bc, storage=P.STORE_F64)
    lp = P.Domain(ctx, dims, cl, 0.8, bc, storage=storage)
    ref.step(2000)
    lp.step(2000)
    _, u64 = ref.macros()
    _, ulp = lp.macros()
    err = np.linalg.norm(ulp - u64) / np.linalg.norm(u64)
    print("storage %d velocity L2 rel error %.3e" % (storage, err), flush=True)
    assert err < bound


def test_mass_conserved_closed_box_clean(P, oracle, ctx):
    dims = (20, 18, 16)
    flags = box_flags(dims, (1, 1, 1), (19, 17, 15))
    g = oracle.make_grid((0, 0, 0), 1.0, dims, (0, 19, 0, 17, 0, 15))
    cl = oracle.classify(g, flags, [])
    for storage, tol in ((0, 1e-12), (1, 2e-6)):
        dom = P.Domain(ctx, dims, cl, 0.7, [], storage=storage, quirks=0)
        dom.set_f(near_eq_state(oracle, 19, cl, np.random.default_rng(9)))
        m0, _ = dom.reduce()
        dom.step(200)
        m1, _ = dom.reduce()
        assert abs(m1 - m0) < tol * m0
