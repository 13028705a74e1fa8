"""GPU checks of the parts whose oracle functions are pinned by closed forms:
the reference's PV2013 inlet waveform, config 1's analytic Poiseuille flow,
the wall shear stress of a linear velocity field, state replacement in the
middle of an AA run, and the batch builder's argument checks."""
import numpy as np
import pytest

from lbm_cases import assert_poiseuille, pipe_case, poiseuille_metrics
from test_gpu_lbm import bits_equal, near_eq_state, six_face_box
from test_oracle_pins import equilibrium_of, walled_box

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


@pytest.mark.parametrize("scheme,options", [(0, 0), (1, 0), (1, 2)])
def test_pv2013_waveform_within_1e12(P, oracle, ctx, scheme, options):
    """The reference's piecewise-polynomial inlet (LBM_WAVE_PV2013, period 0.9 s,
    dt chosen so that 400 steps cross two periods and the zero gap) against the
    oracle: pow/floor/fmod on the GPU and in glibc may differ in the last bit,
    so the moments are compared to 1e-12.  Graph replay and eager steps."""
    dims = (16, 16, 40)
    cl, bcs = pipe_case(oracle, dims, 6.0, 0.08, waveform=0)
    print("PV2013 waveform scheme %d options %d" % (scheme, options), flush=True)
    dom = P.Domain(ctx, dims, cl, 0.8, [P.make_bc(**b) for b in bcs], scheme=scheme, options=options)
    ref = oracle.Sim(19, dims, cl, 0.8, [oracle.make_bc(**b) for b in bcs], quirks=1)
    dt = 0.9 / 150.0
    dom.set_waveform(0, P.WAVE_PV2013, 0.9, dt=dt)
    ref.set_waveform(0, 2, 0.9, dt=dt)
    for n in (1, 140, 259):
        dom.step(n)
        ref.step(n)
        r1, u1 = dom.macros()
        r2, u2 = ref.macros()
        assert np.abs(r1 - r2).max() <= 1e-12 * np.abs(r2).max(), ref.iter
        assert np.linalg.norm(u1 - u2) <= 1e-12 * np.linalg.norm(u2), ref.iter
    assert ref.iter == 400


def test_poiseuille_config1_clean_on_gpu(P, oracle, ctx):
    """Config 1 (32x32x128, R 12, tau 0.8, u_max 0.05, 2000 steps) with clean
    bounce-back: bit-identical to the oracle, and the analytic Poiseuille
    profile and Hagen-Poiseuille relation hold on the GPU result."""
    dims, R, umax, tau = (32, 32, 128), 12.0, 0.05, 0.8
    cl, bcs = pipe_case(oracle, dims, R, umax)
    print("config 1 pipe, clean bounce-back, 2000 steps", flush=True)
    dom = P.Domain(ctx, dims, cl, tau, [P.make_bc(**b) for b in bcs], quirks=0, scheme=P.SCHEME_AA)
    dom.step(2000)
    rho, u = dom.macros()
    m = poiseuille_metrics(dims, R, tau, cl, rho, u)
    print("sections", m["sections"], "flux spread %.2e" % m["flux_spread"], flush=True)
    assert_poiseuille(m, R, umax)
    ref = oracle.Sim(19, dims, cl, tau, [oracle.make_bc(**b) for b in bcs], quirks=0)
    ref.step_threads(2000, min(16, oracle.threads_available()))
    assert bits_equal(dom.f(), ref.f())


@pytest.mark.parametrize("q", [19, 27])
def test_wss_of_a_linear_velocity_field_on_gpu(P, oracle, ctx, q):
    """lbm_get_wss on u = u0 + G x at equilibrium: mu |G m| at every wall node,
    m the mean of c_d over the links whose node exists (rest link included)."""
    dims, tau = (7, 8, 6), 0.74
    flags = walled_box(dims)
    c, w, _ = oracle.lattice(q)
    nx, ny, nz = dims
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    X = np.stack([i.reshape(-1), j.reshape(-1), k.reshape(-1)]).astype(float)
    rng = np.random.default_rng(q)
    G = 1e-3 * rng.standard_normal((3, 3))
    u = 0.01 * rng.standard_normal(3)[:, None] + G @ X
    dom = P.Domain(ctx, dims, flags, tau, [], q=q)
    dom.set_f(equilibrium_of(c, w, u))
    got = dom.wss()
    mu = (2 * tau - 1) / 6
    for st in np.flatnonzero(flags == 2):
        x = X[:, st].astype(int)
        dst = x[None, :] + c
        present = np.all((dst >= 0) & (dst < np.array(dims)), axis=1)
        exp = mu * np.linalg.norm(G @ (c[present].sum(0) / present.sum()))
        assert abs(got[st] - exp) <= 1e-11 * exp + 1e-18, (x, got[st], exp)
    assert np.all(got[flags != 2] == 0.0)


@pytest.mark.parametrize("storage", [0, 1])
def test_aa_set_f_after_an_odd_step(P, oracle, ctx, storage):
    """lbm_set_f re-initialises the AA lattice in the layout a gather step
    reads, whatever the iteration: after an odd number of steps the next step
    must again be a gather.  FP64 against the oracle bit for bit; FP32 against
    the two-lattice scheme bit for bit.  Graph replays (64 steps) included."""
    dims = (10, 11, 12)
    cl, bcs = six_face_box(oracle, dims)
    bc = [P.make_bc(**b) for b in bcs]
    f0 = near_eq_state(oracle, 19, cl, np.random.default_rng(21))
    aa = P.Domain(ctx, dims, cl, 0.7, bc, storage=storage, scheme=P.SCHEME_AA)
    two = P.Domain(ctx, dims, cl, 0.7, bc, storage=storage, scheme=P.SCHEME_TWO_LATTICE)
    ref = oracle.Sim(19, dims, cl, 0.7, [oracle.make_bc(**b) for b in bcs], quirks=1)
    print("AA set_f after odd steps, storage", storage, flush=True)
    for first, after in ((1, 70), (3, 65), (2, 9)):
        for d in (aa, two):
            d.step(first)
        ref.step(first)
        for d in (aa, two):
            d.set_f(f0)
        ref.set_f(f0)
        for d in (aa, two):
            d.step(after)
        ref.step(after)
        assert aa.iteration == ref.iter
        got = aa.f()
        assert bits_equal(got, two.f()), "after %d + %d steps" % (first, after)
        if storage == 0:
            assert bits_equal(got, ref.f()), "after %d + %d steps" % (first, after)


def test_batch_rejects_mixed_or_periodic_cases(P, oracle, ctx):
    """lbm_batch_create: periodic axes would wrap at the packed box's faces and
    link cases to each other, so they are rejected; cases must also agree on
    options, node order and block size."""
    dims = (12, 12, 20)
    cl, bcs = pipe_case(oracle, dims, 4.5, 0.02)
    bc = [P.make_bc(**b) for b in bcs]
    base = dict(dims=dims, flags=cl, tau=0.7, bcs=bc, q=19, storage=P.STORE_F64)

    def expect(cases, text):
        with pytest.raises(P.capi.LbmError) as e:
            P.batch(ctx, cases)
        assert "(-22)" in str(e.value) and text in str(e.value), str(e.value)

    print("batch argument checks", flush=True)
    expect([base, dict(base, options=P.capi.OPT_PERIODIC_X)], "options/fast_axis/block")
    per = P.capi.OPT_PERIODIC_Z
    expect([dict(base, options=per), dict(base, options=per)], "periodic")
    expect([base, dict(base, block=128)], "options/fast_axis/block")
    expect([base, dict(base, fast_axis=0)], "options/fast_axis/block")
    dom, info = P.batch(ctx, [base, dict(base)])
    dom.step(3)
    assert dom.iteration == 3
