"""Time-loop checks of the oracle (Simulation.cpp:903-932) that pin it to
properties the reference's own step satisfies."""
import numpy as np
import pytest

from lbm_cases import box_flags, pipe_case


def closed_box(oracle, dims):
    flags = box_flags(dims, (1, 1, 1), tuple(d - 1 for d in dims))
    g = oracle.make_grid((0, 0, 0), 1.0, dims, (0, dims[0] - 1, 0, dims[1] - 1, 0, dims[2] - 1))
    return oracle.classify(g, flags, [])


@pytest.mark.parametrize("q", [19, 27])
def test_pure_streaming_shift(oracle, q):
    """tau -> inf makes collision the identity; a tagged interior population
    then moves exactly one lattice link per step (push == pull)."""
    dims = (12, 11, 10)
    flags = closed_box(oracle, dims)
    c, w, _ = oracle.lattice(q)
    s = oracle.Sim(q, dims, flags, 1e300, [], quirks=1)
    f = s.f()
    x0 = np.array([5, 5, 5])
    site = lambda p: p[0] + dims[0] * (p[1] + dims[1] * p[2])
    for d in range(q):
        f[d, site(x0)] += 0.125 * (d + 1)
    s.set_f(f)
    s.step(2)
    g = s.f()
    for d in range(q):
        assert g[d, site(x0 + 2 * c[d])] == w[d] + 0.125 * (d + 1)


@pytest.mark.parametrize("q", [19, 27])
def test_rest_state_is_fixed(oracle, q):
    dims = (9, 8, 7)
    flags = closed_box(oracle, dims)
    _, w, _ = oracle.lattice(q)
    s = oracle.Sim(q, dims, flags, 0.8, [], quirks=1)
    s.step(50)
    f = s.f()
    fl = flags > 0
    assert np.allclose(f[:, fl], w[:, None], rtol=2e-15, atol=0)


@pytest.mark.parametrize("q", [19, 27])
def test_mass_conserved_closed_box_clean(oracle, q):
    dims = (10, 9, 8)
    flags = closed_box(oracle, dims)
    c, w, _ = oracle.lattice(q)
    s = oracle.Sim(q, dims, flags, 0.7, [], quirks=0)
    rng = np.random.default_rng(1)
    f = s.f()
    fl = flags > 0
    f[:, fl] = w[:, None] * (1 + 0.02 * rng.standard_normal((q, fl.sum())))
    s.set_f(f)
    m0 = s.mass()
    s.step(40)
    assert abs(s.mass() - m0) < 1e-12 * m0


def test_poiseuille_pipe(oracle):
    """Steady pipe flow (clean bounce-back) is a parabola in r with a flux that
    is the same at every cross-section (D3Q19).  The reference-quirk mode leaks
    mass through the +x wall links it does not bounce back, so the check runs
    on the clean mode."""
    dims, R, umax = (14, 14, 40), 5.5, 0.02
    cl, bcs = pipe_case(oracle, dims, R, umax)
    s = oracle.Sim(19, dims, cl, 0.8, [oracle.make_bc(**b) for b in bcs], quirks=0)
    s.step(900)
    rho, u = s.macros()
    nx, ny, nz = dims
    uz = u[2].reshape(nz, ny, nx)
    rr = rho.reshape(nz, ny, nx)
    flux = [(rr[k] * uz[k]).sum() for k in (nz // 4, nz // 2, 3 * nz // 4)]
    assert max(flux) - min(flux) < 5e-3 * max(flux)
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    r2 = ((i - (nx - 1) / 2) ** 2 + (j - (ny - 1) / 2) ** 2).astype(float)
    inner = r2 < (R - 1.0) ** 2
    prof = uz[nz // 2][inner]
    A = np.stack([np.ones(inner.sum()), r2[inner]], 1)
    (a, b), *_ = np.linalg.lstsq(A, prof, rcond=None)
    assert np.linalg.norm(A @ [a, b] - prof) < 0.02 * np.linalg.norm(prof)
    assert R - 1.0 < np.sqrt(-a / b) < R + 1.0


def shear_wave(oracle, q, dims, tau, amp=1e-3):
    """Fully periodic box with u_x = amp sin(2 pi y / ny) at equilibrium."""
    c, w, _ = oracle.lattice(q)
    nx, ny, nz = dims
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ux = (amp * np.sin(2 * np.pi * j / ny)).reshape(-1)
    cu = c[:, 0][:, None] * ux[None, :]
    f = w[:, None] * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * ux * ux)
    return np.ones(nx * ny * nz, np.uint8), f, ux


@pytest.mark.parametrize("q", [19, 27])
def test_periodic_shear_wave_decays_at_the_lattice_viscosity(oracle, q):
    """Closed form: u_x(y, t) = U sin(k y) exp(-nu k^2 t), nu = (tau - 1/2)/3."""
    dims, tau, steps = (4, 32, 4), 0.8, 200
    flags, f0, ux0 = shear_wave(oracle, q, dims, tau)
    s = oracle.Sim(q, dims, flags, tau, [], quirks=1, periodic=7)
    s.set_f(f0)
    s.step(steps)
    _, u = s.macros()
    kk = 2 * np.pi / dims[1]
    nu = (tau - 0.5) / 3.0
    ratio = (u[0] @ ux0) / (ux0 @ ux0)
    assert abs(ratio - np.exp(-nu * kk * kk * steps)) < 2e-3
    assert np.abs(u[1:]).max() < 1e-12


@pytest.mark.parametrize("q,quirks", [(19, 1), (19, 0), (27, 1)])
def test_threaded_step_equals_sequential(oracle, q, quirks):
    """The OpenMP step (CPU baseline) gives the bits of the sequential step:
    pulsatile inlet, pressure outlet, walls, both quirk modes."""
    dims = (14, 14, 30)
    cl, bcs = pipe_case(oracle, dims, 5.5, 0.04, waveform=0)
    a, phi = np.array([0.3, 0.1]), np.array([0.2, 1.3])
    sims = []
    for _ in range(2):
        s = oracle.Sim(q, dims, cl, 0.7, [oracle.make_bc(**b) for b in bcs], quirks=quirks)
        s.set_waveform(0, 1, 50.0, a, phi)
        sims.append(s)
    sims[0].step(40)
    sims[1].step_threads(40, 4)
    assert np.array_equal(sims[0].f(), sims[1].f())


def test_poiseuille_config1_clean(oracle):
    """Config 1 (32x32x128 z-pipe, R 12, tau 0.8, parabolic inlet u_max 0.05,
    outlet rho 1.0, 2000 steps from rest) with clean bounce-back: the analytic
    Poiseuille profile and the Hagen-Poiseuille pressure-gradient relation."""
    from lbm_cases import assert_poiseuille, poiseuille_metrics
    dims, R, umax, tau = (32, 32, 128), 12.0, 0.05, 0.8
    cl, bcs = pipe_case(oracle, dims, R, umax)
    s = oracle.Sim(19, dims, cl, tau, [oracle.make_bc(**b) for b in bcs], quirks=0)
    s.step_threads(2000, min(8, oracle.threads_available()))
    rho, u = s.macros()
    assert_poiseuille(poiseuille_metrics(dims, R, tau, cl, rho, u), R, umax)
