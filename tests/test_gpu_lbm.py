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


def both(P, oracle, ctx, q, dims, flags, tau, bcs, quirks=1, storage=0, fast_axis=-1, options=0):
    dom = P.Domain(ctx, dims, flags, tau, [P.make_bc(**b) for b in bcs], q=q, storage=storage,
                   quirks=quirks, fast_axis=fast_axis, options=options)
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


# step kernels: row chunks in one launch (two-lattice: chunk pairs), records +
# segments in two launches, generic per-node kernel, one row chunk per wavefront
LAYOUTS = [0, 64, 32, 256]


@pytest.mark.parametrize("q", [19, 27])
@pytest.mark.parametrize("quirks", [1, 0])
@pytest.mark.parametrize("options", LAYOUTS)
def test_fp64_bitwise_all_faces(P, oracle, ctx, q, quirks, options):
    dims = (10, 11, 12)
    cl, bcs = six_face_box(oracle, dims)
    print("six-face box q=%d quirks=%d options=%d" % (q, quirks, options), flush=True)
    dom, ref = both(P, oracle, ctx, q, dims, cl, 0.7, bcs, quirks=quirks, options=options)
    st = dom.stats()
    assert st["layout"] == {0: 1, 64: 2, 32: 0, 256: 1}[options]
    if options == 0:  # the words cut by row ends hold chunk pairs
        assert st["step_waves"] < st["segments"]
    if options == 256:
        assert st["step_waves"] == st["segments"]
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


@pytest.mark.parametrize("fast_axis,options", [(-1, 0), (0, 0), (1, 0), (-1, 1), (-1, 2), (-1, 32), (-1, 64),
                                               (0, 32), (-1, 256)])
def test_fp64_bitwise_pipe_config1(P, oracle, ctx, fast_axis, options):
    """D3Q19 FP64, 32x32x128 z-pipe of radius 12, tau 0.8, parabolic inlet
    u_max 0.05, outlet rho 1.0, 2000 steps from rest (the small plain case);
    shorter runs for the other node orders and with the row fast path off."""
    dims = (32, 32, 128)
    cl, bcs = pipe_case(oracle, dims, 12.0, 0.05)
    steps = 2000 if (fast_axis, options) == (-1, 0) else 200
    print("pipe 32x32x128 fast_axis=%d options=%d %d steps" % (fast_axis, options, steps), flush=True)
    dom, ref = both(P, oracle, ctx, 19, dims, cl, 0.8, bcs, fast_axis=fast_axis, options=options)
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


@pytest.mark.parametrize("storage,bound", [(1, 2e-6), (2, 5e-3)])
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
    assert err < bound, "storage %d: velocity L2 error %.3e vs FP64 exceeds %.1e" % (storage, err, bound)


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


def test_batch_matches_individual_runs(P, oracle, ctx):
    """Three pipes of different sizes, one with a Fourier inlet, packed into one
    domain: every case is bit-identical to the same case run alone."""
    specs = [((14, 14, 40), 5.5, 0.03, -1), ((20, 18, 33), 7.5, 0.02, 0), ((12, 12, 57), 4.5, 0.04, -1)]
    cases, singles = [], []
    rng = np.random.default_rng(5)
    a, phi = 0.3 * rng.random(4), 2 * np.pi * rng.random(4)
    for dims, R, um, wf in specs:
        cl, bcs = pipe_case(oracle, dims, R, um, waveform=wf)
        pb = [P.make_bc(**b) for b in bcs]
        cases.append(dict(dims=dims, flags=cl, tau=0.7, bcs=pb, q=19, storage=P.STORE_F64))
        d = P.Domain(ctx, dims, cl, 0.7, pb)
        if wf >= 0:
            d.set_waveform(wf, P.WAVE_FOURIER, 300.0, a, phi)
        singles.append(d)
    dom, info = P.batch(ctx, cases)
    dom.set_waveform(info[1]["wave_base"] + 0, P.WAVE_FOURIER, 300.0, a, phi)
    dom.step(150)
    for d, inf in zip(singles, info):
        d.step(150)
        r1, u1 = d.macros()
        r2, u2 = dom.macros_box(inf["origin"], inf["dims"])
        assert bits_equal(r1, r2) and bits_equal(u1, u2)
        assert bits_equal(d.wss(), dom.wss_box(inf["origin"], inf["dims"]))


@pytest.mark.parametrize("storage", [0, 1])
def test_waveform_ensemble_on_one_geometry(P, oracle, ctx, storage):
    """An ensemble of inlet waveforms: four copies of one pipe, each with its
    own seeded Fourier waveform, packed into one domain (one launch per step,
    AA): each copy equals the same pipe run alone with its waveform, bit for
    bit, in FP64 and in FP32 storage."""
    dims, R, um = (14, 15, 36), 5.5, 0.04
    cl, bcs = pipe_case(oracle, dims, R, um, waveform=0)
    pb = [P.make_bc(**b) for b in bcs]
    rng = np.random.default_rng(17)
    waves = [(0.4 * rng.random(3), 2 * np.pi * rng.random(3), 120.0 + 40.0 * k) for k in range(4)]
    cases = [dict(dims=dims, flags=cl, tau=0.65, bcs=pb, q=19, storage=storage, scheme=P.SCHEME_AA)
             for _ in waves]
    dom, info = P.batch(ctx, cases)
    singles = []
    for (a, phi, T), inf in zip(waves, info):
        dom.set_waveform(inf["wave_base"], P.WAVE_FOURIER, T, a, phi)
        d = P.Domain(ctx, dims, cl, 0.65, pb, storage=storage, scheme=P.SCHEME_AA)
        d.set_waveform(0, P.WAVE_FOURIER, T, a, phi)
        singles.append(d)
    dom.step(131)
    us = []
    for d, inf in zip(singles, info):
        d.step(131)
        r1, u1 = d.macros()
        r2, u2 = dom.macros_box(inf["origin"], inf["dims"])
        assert bits_equal(r1, r2) and bits_equal(u1, u2)
        us.append(u1)
        d.close()
    # the waveforms differ, so the copies do
    assert not np.array_equal(us[0], us[1])
    dom.close()


@pytest.mark.parametrize("scheme", [0, 1])
def test_step_timed_equals_step(P, oracle, ctx, scheme):
    """lbm_step_timed builds a missing step graph before its first event; the
    steps it then runs equal lbm_step's bit for bit, from a fresh domain
    (graph built inside the call) and after an odd number of steps (AA starts
    its graphs on even iterations)."""
    dims = (14, 14, 30)
    cl, bcs = pipe_case(oracle, dims, 5.5, 0.04)
    doms = [P.Domain(ctx, dims, cl, 0.8, [P.make_bc(**b) for b in bcs], scheme=scheme) for _ in range(2)]
    for n in (100, 67, 130):
        ms = doms[0].step_timed(n)
        doms[1].step(n)
        assert ms > 0.0
        assert bits_equal(doms[0].f(), doms[1].f()), n
    for d in doms:
        d.close()


@pytest.mark.parametrize("storage", [0, 1, 2])
def test_aa_equals_two_lattice(P, oracle, ctx, storage):
    """The AA scheme (one lattice, alternating gather / local steps) computes
    exactly what the two-lattice pull computes, in every precision; with the
    graph path, the eager path, odd and even step counts.  The two-lattice pull
    with one row chunk per wavefront equals its default (chunk pairs), and
    launches alternating direction (the default; at 64 threads per block the
    box spans several blocks per XCD) equal launches that all walk upward."""
    dims = (16, 16, 40)
    cl, bcs = pipe_case(oracle, dims, 6.0, 0.03, waveform=0)
    bc = [P.make_bc(**b) for b in bcs]
    a, phi = np.array([0.2, 0.1]), np.array([0.3, 1.1])
    two = P.Domain(ctx, dims, cl, 0.7, bc, storage=storage)
    aa = P.Domain(ctx, dims, cl, 0.7, bc, storage=storage, scheme=P.SCHEME_AA)
    aa_eager = P.Domain(ctx, dims, cl, 0.7, bc, storage=storage, scheme=P.SCHEME_AA, options=P.capi.OPT_NO_GRAPH)
    aa_generic = P.Domain(ctx, dims, cl, 0.7, bc, storage=storage, scheme=P.SCHEME_AA,
                          options=P.capi.OPT_GENERIC_KERNEL)
    aa_split = P.Domain(ctx, dims, cl, 0.7, bc, storage=storage, scheme=P.SCHEME_AA, options=P.capi.OPT_SPLIT_LAUNCH)
    two_single = P.Domain(ctx, dims, cl, 0.7, bc, storage=storage, options=P.capi.OPT_ONE_CHUNK_PER_WAVE)
    aa64 = P.Domain(ctx, dims, cl, 0.7, bc, storage=storage, scheme=P.SCHEME_AA, block=64)
    aa64_fwd = P.Domain(ctx, dims, cl, 0.7, bc, storage=storage, scheme=P.SCHEME_AA, block=64,
                        options=P.capi.OPT_FORWARD_ONLY)
    two64 = P.Domain(ctx, dims, cl, 0.7, bc, storage=storage, block=64)
    two64_fwd = P.Domain(ctx, dims, cl, 0.7, bc, storage=storage, block=64, options=P.capi.OPT_FORWARD_ONLY)
    others = (aa, aa_eager, aa_generic, aa_split, two_single, aa64, aa64_fwd, two64, two64_fwd)
    for d in (two,) + others:
        d.set_waveform(0, P.WAVE_FOURIER, 200.0, a, phi)
    for n in (1, 70, 3, 130):
        for d in (two,) + others:
            d.step(n)
        f2 = two.f()
        for k, d in enumerate(others):
            assert bits_equal(d.f(), f2), "AA variant %d after %d" % (k, two.iteration)
        r2, u2 = two.macros()
        r1, u1 = aa.macros()
        assert bits_equal(r1, r2) and bits_equal(u1, u2)


@pytest.mark.parametrize("q", [19, 27])
@pytest.mark.parametrize("scheme", [0, 1])
@pytest.mark.parametrize("case", ["six_faces", "pipe"])
def test_fp16_record_paths_identical(P, oracle, ctx, q, scheme, case):
    """FP16 row chunks that hold boundary records: the path through one 32-bit
    descriptor over the whole lattice (the default below 4 GiB) and the path
    through 64-bit addresses (LBM_OPT_WIDE_RECORDS, automatic above) give the
    same bits.  The six-face box has a Zou-He condition on every face and stale
    populations; the pipe has bounce-back walls, a profiled inlet and an outlet."""
    if case == "six_faces":
        dims = (10, 11, 12)
        cl, bcs = six_face_box(oracle, dims)
    else:
        dims = (14, 14, 30)
        cl, bcs = pipe_case(oracle, dims, 5.5, 0.03)
    bc = [P.make_bc(**b) for b in bcs]
    f0 = near_eq_state(oracle, q, cl, np.random.default_rng(5))
    doms = [P.Domain(ctx, dims, cl, 0.7, bc, q=q, storage=2, scheme=scheme, options=o)
            for o in (0, P.capi.OPT_WIDE_RECORDS)]
    for d in doms:
        d.set_f(f0)
    for n in (1, 2, 9):
        for d in doms:
            d.step(n)
        assert bits_equal(doms[0].f(), doms[1].f()), "after %d steps" % doms[0].iteration
    r0, u0 = doms[0].macros()
    r1, u1 = doms[1].macros()
    assert bits_equal(r0, r1) and bits_equal(u0, u1)


def test_aa_fp64_matches_oracle_six_faces(P, oracle, ctx):
    dims = (10, 11, 12)
    cl, bcs = six_face_box(oracle, dims)
    dom = P.Domain(ctx, dims, cl, 0.7, [P.make_bc(**b) for b in bcs], q=27, scheme=P.SCHEME_AA)
    ref = oracle.Sim(27, dims, cl, 0.7, [oracle.make_bc(**b) for b in bcs], quirks=1)
    f0 = near_eq_state(oracle, 27, cl, np.random.default_rng(11))
    dom.set_f(f0)
    ref.set_f(f0)
    for n in (1, 2, 17):
        dom.step(n)
        ref.step(n)
        assert bits_equal(dom.f(), ref.f())
    assert bits_equal(dom.wss(), ref.wss())


@pytest.mark.parametrize("q", [19, 27])
@pytest.mark.parametrize("scheme", [0, 1])
def test_periodic_shear_wave(P, oracle, ctx, q, scheme):
    """Periodic axes (ghost-filled site map) against the oracle's wrapped
    neighbours: FP64 bitwise, FP32 to 1e-5 of the amplitude; the wave decays at nu = (tau-1/2)/3."""
    from test_oracle_lbm import shear_wave
    dims, tau, steps = (6, 32, 5), 0.8, 150
    flags, f0, ux0 = shear_wave(oracle, q, dims, tau)
    per = P.capi.OPT_PERIODIC_X | P.capi.OPT_PERIODIC_Y | P.capi.OPT_PERIODIC_Z
    ref = oracle.Sim(q, dims, flags, tau, [], quirks=1, periodic=7)
    ref.set_f(f0)
    ref.step(steps)
    for storage, kern in ((0, 0), (0, P.capi.OPT_GENERIC_KERNEL), (1, 0)):
        dom = P.Domain(ctx, dims, flags, tau, [], q=q, storage=storage, scheme=scheme, options=per | kern)
        dom.set_f(f0)
        dom.step(steps)
        if storage == 0:
            got = dom.f()
            if not bits_equal(got, ref.f()):
                c, _, _ = oracle.lattice(q)
                ug = (c[:, 0] @ got) / got.sum(0)
                pytest.fail("periodic scheme %d q %d: amplitude ratio to the initial wave %.6f, oracle %.6f"
                            % (scheme, q, (ug @ ux0) / (ux0 @ ux0), (ref.macros()[1][0] @ ux0) / (ux0 @ ux0)))
        else:
            _, u = dom.macros()
            _, ur = ref.macros()
            assert np.abs(u - ur).max() < 1e-5 * np.abs(ur).max()
    _, u = ref.macros()
    kk = 2 * np.pi / dims[1]
    assert abs((u[0] @ ux0) / (ux0 @ ux0) - np.exp(-(tau - 0.5) / 3 * kk * kk * steps)) < 2e-3


def test_autotune_picks_a_valid_layout_and_caches_it(P, oracle, ctx, tmp_path):
    """lbm_autotune times every candidate (scheme x node order x kernel
    layout, then block sizes for the winner), returns one that builds and
    reproduces the FP64 oracle bitwise, and answers from its cache file the
    second time."""
    dims = (16, 16, 40)
    cl, bcs = pipe_case(oracle, dims, 6.0, 0.03)
    cache = str(tmp_path / "tune.txt")
    kw = dict(dims=dims, flags=cl, tau=0.8, bcs=[P.make_bc(**b) for b in bcs], q=19, storage=P.STORE_F64)
    r1 = P.autotune(ctx, steps=20, cache_path=cache, **kw)
    assert r1["ntried"] == 2 * 3 * 3 + 3 and r1["mlups"] > 0  # then the three other block sizes
    assert r1["block"] in (64, 128, 256, 512) and r1["scheme"] in (0, 1) and r1["fast_axis"] in (0, 1, 2)
    r2 = P.autotune(ctx, steps=20, cache_path=cache, **kw)
    assert r2["ntried"] == 0 and r2["key"] == r1["key"]
    assert (r2["scheme"], r2["fast_axis"], r2["options"], r2["block"]) == \
        (r1["scheme"], r1["fast_axis"], r1["options"], r1["block"])
    dom = P.Domain(ctx, dims, cl, 0.8, kw["bcs"], scheme=r1["scheme"], fast_axis=r1["fast_axis"],
                   options=r1["options"], block=r1["block"])
    ref = oracle.Sim(19, dims, cl, 0.8, [oracle.make_bc(**b) for b in bcs], quirks=1)
    dom.step(130)
    ref.step(130)
    assert bits_equal(dom.f(), ref.f())


@pytest.mark.parametrize("block", [64, 128, 256, 512])
def test_block_sizes_give_identical_bits(P, oracle, ctx, block):
    dims = (10, 11, 12)
    cl, bcs = six_face_box(oracle, dims)
    for options in LAYOUTS:
        dom, ref = both(P, oracle, ctx, 19, dims, cl, 0.7, bcs, options=options)
        alt = P.Domain(ctx, dims, cl, 0.7, [P.make_bc(**b) for b in bcs], q=19, options=options, block=block)
        f0 = near_eq_state(oracle, 19, cl, np.random.default_rng(3))
        dom.set_f(f0)
        alt.set_f(f0)
        dom.step(9)
        alt.step(9)
        assert bits_equal(dom.f(), alt.f()), "block %d options %d" % (block, options)
