"""Pins of the oracle functions that no reference fixture covers: the inlet
waveforms (pulsatile.h:12-28, Simulation.cpp:875-900), the wall shear stress
(Simulation.cpp:1374-1421) and the reference's step quirks (Simulation.cpp:
1353-1370 bounce-back from direction 1, stale populations at unlinked links).

Every expected value here comes from an independent evaluation (long-double
arithmetic, a closed form of a linear velocity field, or the node's own
pre-step populations), never from the oracle's own formula."""
import numpy as np
import pytest

from lbm_cases import box_flags, pipe_case

LD = np.longdouble
EPS = np.finfo(np.float64).eps

# pulsatile.h:14-28, highest power first, as decimal literals read in long double
PV_POLY = [
    ["-8.579e+08", "4.059e+08", "-7.817e+07", "7.994e+06", "-4.835e+05", "1.782e+04", "-321.6", "3.215",
     "0.1423"],
    ["-4.32e+06", "7.94e+06", "-6.298e+06", "2.82e+06", "-7.809e+05", "1.37e+05", "-1.484e+04", "901.9",
     "-22.87"],
    ["-5.123", "14.15", "-14.21", "5.865", "-0.5696"],
]


def horner_ld(coef_text, x):
    acc = LD(0)
    for c in coef_text:
        acc = acc * x + LD(c)
    return acc


def abs_terms(coef_text, x):
    """sum |c_k| x^k and sum k |c_k| x^(k-1) in double (error-bound scale)."""
    deg = len(coef_text) - 1
    s = ds = 0.0
    for i, c in enumerate(coef_text):
        k = deg - i
        s += abs(float(c)) * x ** k
        if k:
            ds += k * abs(float(c)) * x ** (k - 1)
    return s, ds


def one_node_sim(oracle):
    return oracle.Sim(19, (2, 2, 2), np.ones(8, np.uint8), 0.8, [], quirks=1)


def wave(oracle, sim, it):
    return oracle.lib().orc_waveform_eval(sim.h, 0, int(it))


@pytest.mark.parametrize("dt", [0.9 / 157.3, 1e-3, 2.7e-4])
def test_pv2013_waveform_against_long_double(oracle, dt):
    """The reference's piecewise polynomial over a 0.9 s period, reduced to the
    period, with the zero gap in [0.81, 0.9): evaluated in long double by
    Horner's rule and compared within a rounding bound (the double result has
    at most a few roundings per term plus the reduction's error times the
    slope).  Samples within 1e-12 of a segment or period boundary are skipped,
    since there the segment choice depends on the last bit of the time."""
    s = one_node_sim(oracle)
    period = 0.9
    s.set_waveform(0, 2, period, dt=dt)
    P = LD(period)
    edges = [LD(0.0), LD(0.1 * period), LD(0.3 * period), LD(0.9 * period), P]
    checked = gap = skipped = 0
    for it in range(0, 9000, 7):
        x = LD(float(it) * dt)  # the oracle's time, t = it * dt in double
        r = x - P * np.floor(x / P)
        if min(abs(r - e) for e in edges) < 1e-12:
            skipped += 1
            continue
        got = wave(oracle, s, it)
        seg = 0 if r < edges[1] else (1 if r < edges[2] else (2 if r < edges[3] else 3))
        if seg == 3:
            assert got == 0.0, "t=%r in the zero gap" % float(x)
            gap += 1
            continue
        exp = horner_ld(PV_POLY[seg], r)
        scale, slope = abs_terms(PV_POLY[seg], float(r))
        bound = 32 * EPS * scale + slope * 8 * EPS * float(x) + 1e-300
        assert abs(LD(got) - exp) <= bound, (it, float(x), got, float(exp))
        checked += 1
    assert checked > 800 and gap > 50 and skipped < 10


def test_pv2013_waveform_first_call_and_period(oracle):
    """Iteration 0 evaluates the first polynomial at 0 without the period
    reduction (Simulation.cpp:876-883: the first call skips it): exactly the
    constant term 0.1423.  The value at an iteration does not depend on which
    iteration was evaluated first, and W(t + period) = W(t) to rounding."""
    a = one_node_sim(oracle)
    b = one_node_sim(oracle)
    dt = 0.9 / 200.0
    for s in (a, b):
        s.set_waveform(0, 2, 0.9, dt=dt)
    late = [wave(oracle, b, it) for it in (1234, 57, 0)]
    early = [wave(oracle, a, it) for it in (0, 57, 1234)][::-1]
    assert late == early
    assert wave(oracle, a, 0) == 0.1423
    for it in (3, 41, 77, 133, 171):
        w0, w1 = wave(oracle, a, it), wave(oracle, a, it + 200)
        assert abs(w0 - w1) <= 1e-9 * max(1.0, abs(w0)), it


def test_fourier_waveform_against_long_double(oracle):
    """W(t) = 1 + sum_k a_k sin(2 pi k t / T + phi_k) against long double, and
    the closed form of one mode at a quarter period."""
    s = one_node_sim(oracle)
    rng = np.random.default_rng(7)
    a, phi = 0.3 * rng.random(4), 2 * np.pi * rng.random(4)
    T = 4000.0
    s.set_waveform(0, 1, T, a, phi)
    twopi = LD("6.283185307179586476925286766559005768")
    for it in list(range(0, 200, 3)) + list(range(7000, 8000, 37)):
        exp = LD(1)
        bound = 4 * EPS
        for k in range(4):
            arg = twopi * (k + 1) * it / LD(T) + LD(phi[k])
            exp += LD(a[k]) * np.sin(arg)
            bound += abs(a[k]) * (8 * EPS * float(abs(arg)) + 4 * EPS)
        assert abs(LD(wave(oracle, s, it)) - exp) <= bound, it
    one = one_node_sim(oracle)
    one.set_waveform(0, 1, 400.0, [0.5], [0.0])
    assert abs(wave(oracle, one, 100) - 1.5) < 1e-15
    assert abs(wave(oracle, one, 300) - 0.5) < 1e-15
    assert wave(oracle, one, 0) == 1.0


# ---------------------------------------------------------------- wall shear


def walled_box(dims):
    """Every site a node: type 2 (wall) on the box faces, 1 inside."""
    nx, ny, nz = dims
    f = np.full((nz, ny, nx), 2, np.uint8)
    f[1:-1, 1:-1, 1:-1] = 1
    return f.reshape(-1)


def equilibrium_of(c, w, u):
    cu = c @ u
    usq = (u * u).sum(0)
    return w[:, None] * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq)


@pytest.mark.parametrize("q", [19, 27])
def test_wss_of_a_linear_velocity_field(oracle, q):
    """u(x) = u0 + G x at rho = 1 (populations at equilibrium, whose moments
    are exactly rho and u).  The reference's gradient at a wall node is the mean
    of u(x + c_d) - u(x) = G c_d over the links d whose node exists, the rest
    link included (count |P|), so WSS = mu |G m| with m = sum_P c_d / |P| and
    mu = (2 tau - 1)/6, computed here from the box geometry alone."""
    dims, tau = (7, 8, 6), 0.74
    flags = walled_box(dims)
    c, w, _ = oracle.lattice(q)
    nx, ny, nz = dims
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    X = np.stack([i.reshape(-1), j.reshape(-1), k.reshape(-1)]).astype(float)
    rng = np.random.default_rng(q)
    G = 1e-3 * rng.standard_normal((3, 3))
    u0 = 0.01 * rng.standard_normal(3)
    u = u0[:, None] + G @ X
    s = oracle.Sim(q, dims, flags, tau, [], quirks=1)
    s.set_f(equilibrium_of(c, w, u))
    got = s.wss()
    mu = (2 * tau - 1) / 6
    nwall = 0
    for st in np.flatnonzero(flags == 2):
        x = X[:, st].astype(int)
        dst = x[None, :] + c
        present = np.all((dst >= 0) & (dst < np.array(dims)), axis=1)  # every site is a node
        m = c[present].sum(0) / present.sum()
        exp = mu * np.linalg.norm(G @ m)
        assert abs(got[st] - exp) <= 1e-11 * exp + 1e-18, (x, got[st], exp)
        nwall += 1
    assert nwall == flags.size - (nx - 2) * (ny - 2) * (nz - 2)
    assert np.all(got[flags != 2] == 0.0)


# ------------------------------------------------------------- step quirks


def bgk(c, w, omega, f):
    """BGK post-collision populations of f (nodes in columns), textbook form;
    columns of non-nodes (all zero) stay zero."""
    rho = f.sum(0)
    u = (c.T @ f) / np.where(rho == 0, 1.0, rho)
    feq = rho * equilibrium_of(c, w, u)
    return omega * feq + (1 - omega) * f


def random_state(oracle, q, flags, seed, amp=0.02):
    c, w, _ = oracle.lattice(q)
    rng = np.random.default_rng(seed)
    f = w[:, None] * (1 + amp * rng.standard_normal((q, flags.size)))
    f[:, flags == 0] = 0.0
    return f


def missing_dst(c, dims, flags, st):
    """Directions d whose node x + c_d does not exist."""
    nx, ny, nz = dims
    x = np.array([st % nx, (st // nx) % ny, st // (nx * ny)])
    dst = x[None, :] + c
    inb = np.all((dst >= 0) & (dst < np.array(dims)), axis=1)
    out = np.zeros(len(c), bool)
    for d in range(len(c)):
        if not inb[d]:
            out[d] = True
        else:
            out[d] = flags[dst[d, 0] + nx * (dst[d, 1] + ny * dst[d, 2])] == 0
    return out


@pytest.mark.parametrize("quirks", [1, 0])
def test_bounce_back_links_and_the_direction0_skip(oracle, quirks):
    """One step from a random state on a closed box.  At a wall node x, for
    every link d whose node x + c_d is missing, f[opp d](x) takes the node's own
    post-collision f_new[d](x) (half-way bounce-back, Node.cpp:121-125).  With
    the reference's loop starting at d = 1 (quirks on, D3Q19), the +x link is
    never bounced: f[1](x) keeps its pre-step value bit for bit."""
    q, dims, tau = 19, (7, 6, 5), 0.7
    flags = walled_box(dims)
    c, w, opp = oracle.lattice(q)
    f0 = random_state(oracle, q, flags, 3)
    s = oracle.Sim(q, dims, flags, tau, [], quirks=quirks)
    s.set_f(f0)
    s.step(1)
    f1 = s.f()
    fnew = bgk(c, w, 1.0 / tau, f0)
    skipped = bounced = 0
    for st in np.flatnonzero(flags == 2):
        miss = missing_dst(c, dims, flags, st)
        for d in np.flatnonzero(miss):
            if quirks and d == 0:
                assert f1[1, st] == f0[1, st]
                skipped += 1
            else:
                assert abs(f1[opp[d], st] - fnew[d, st]) < 1e-15, (st, d)
                bounced += 1
    assert bounced > 100 and (skipped > 0) == bool(quirks)


@pytest.mark.parametrize("quirks", [1, 0])
def test_unlinked_populations_at_inlet_and_outlet_nodes(oracle, quirks):
    """Inlet/outlet nodes get no bounce-back in the reference (only wall nodes
    are in WallBoundaryNode_): a population whose source x - c_d is missing and
    which Zou-He does not rewrite (c_axis != sigma) keeps its pre-step value
    bit for bit.  With clean bounce-back (quirks off) it takes f_new[opp d](x)
    instead."""
    q, dims, tau = 19, (14, 14, 20), 0.7
    cl, bcs = pipe_case(oracle, dims, 5.5, 0.03)
    c, w, opp = oracle.lattice(q)
    f0 = random_state(oracle, q, cl, 8)
    s = oracle.Sim(q, dims, cl, tau, [oracle.make_bc(**b) for b in bcs], quirks=quirks)
    s.set_f(f0)
    s.step(1)
    f1 = s.f()
    fnew = bgk(c, w, 1.0 / tau, f0)
    checked = 0
    for st in np.flatnonzero(cl >= 3):
        bc = bcs[cl[st] - 3]
        miss_src = missing_dst(-c, dims, cl, st)  # node at x - c_d missing
        for d in np.flatnonzero(miss_src):
            if c[d, bc["axis"]] == bc["sigma"]:
                continue  # a Zou-He unknown
            if quirks:
                assert f1[d, st] == f0[d, st], (st, d)
            else:
                assert abs(f1[d, st] - fnew[opp[d], st]) < 1e-15, (st, d)
            checked += 1
    assert checked > 50


@pytest.mark.parametrize("q", [19, 27])
def test_boundary_passes_in_any_order_equal_the_per_node_loop(oracle, q):
    """The reference applies its boundary conditions as three passes over
    separate lists (ApplyZouHeBc over the inlet list, ApplyWallBc over the wall
    list, ApplyOutletBc over the outlet lists; Simulation.cpp:908-918).  The
    oracle, like the GPU kernels, applies them in one loop node by node.  With
    the reference's quirks every node is in one list and each update reads and
    writes only its own node, so the passes in the reference's order, and in
    every other order, give the per-node loop's bits.  The pipe's inlet and
    outlet rims sit next to wall nodes (pulsatile inlet, stale links, the
    direction-0 skip for D3Q19), from a random state, over several steps."""
    import itertools
    dims, tau = (12, 13, 18), 0.7
    cl, bcs = pipe_case(oracle, dims, 5.2, 0.03, waveform=0)
    assert (cl == 2).sum() > 100 and (cl == 3).sum() > 20 and (cl == 4).sum() > 20
    f0 = random_state(oracle, q, cl, 11)

    def sim():
        s = oracle.Sim(q, dims, cl, tau, [oracle.make_bc(**b) for b in bcs], quirks=1)
        s.set_waveform(0, 1, 50.0, [0.3, 0.1], [0.2, 1.0])  # Fourier mode
        s.set_f(f0)
        return s

    ref = sim()
    ref.step(4)
    want = ref.f()
    for order in itertools.permutations(range(3)):
        s = sim()
        s.step_passes(4, order)
        got = s.f()
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), order
