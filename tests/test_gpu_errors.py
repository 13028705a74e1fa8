"""Error paths of the C ABI on a device: bad arguments come back as negative
codes with a message (lbm_last_error), never as a launch, and the context
stays usable afterwards.  Degenerate inputs the reference accepts (no fluid
node, a facet-free mesh, a step count of zero) have a defined outcome."""
import numpy as np
import pytest

from lbm_cases import pipe_case

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


def expect(P, fn, code, text):
    with pytest.raises(P.capi.LbmError) as e:
        fn()
    msg = str(e.value)
    assert "(%d)" % code in msg and text in msg, msg


def test_domain_arguments_rejected(P, oracle, ctx):
    dims = (8, 8, 12)
    cl, bcs = pipe_case(oracle, dims, 3.0, 0.02)
    bc = [P.make_bc(**b) for b in bcs]
    EINVAL = -22

    def make(**kw):
        args = dict(q=19, storage=0)
        args.update(kw)
        tau = args.pop("tau", 0.8)
        flags = args.pop("flags", cl)
        bcl = args.pop("bcs", bc)
        dd = args.pop("dims", dims)
        return P.Domain(ctx, dd, flags, tau, bcl, **args)

    print("domain argument checks", flush=True)
    expect(P, lambda: make(q=15), EINVAL, "q must be 19 or 27")
    expect(P, lambda: make(storage=7), EINVAL, "bad storage")
    expect(P, lambda: make(scheme=5), EINVAL, "scheme")
    expect(P, lambda: make(tau=0.5), EINVAL, "tau must exceed 0.5")
    expect(P, lambda: make(block=96), EINVAL, "block must be")
    expect(P, lambda: make(bcs=bc[:1]), EINVAL, "missing boundary condition")
    expect(P, lambda: make(flags=np.zeros_like(cl)), EINVAL, "no fluid nodes")
    bad = dict(bcs[1])
    bad["rho"] = -1.0
    expect(P, lambda: make(bcs=[bc[0], P.make_bc(**bad)]), EINVAL, "bad boundary condition 1")
    # the context still works
    dom = make()
    dom.step(3)
    assert dom.iteration == 3
    dom.close()


def test_step_and_readback_arguments(P, oracle, ctx):
    dims = (8, 8, 12)
    cl, bcs = pipe_case(oracle, dims, 3.0, 0.02)
    for scheme in (0, 1):
        dom = P.Domain(ctx, dims, cl, 0.8, [P.make_bc(**b) for b in bcs], storage=1, scheme=scheme)
        print("step / readback checks, scheme", scheme, flush=True)
        f0 = dom.f()
        dom.step(0)  # no step: same state, same iteration
        assert dom.iteration == 0 and np.array_equal(dom.f(), f0)
        expect(P, lambda: dom.step(-1), -22, "negative step count")
        expect(P, lambda: dom.set_waveform(9999, P.WAVE_FOURIER, 100.0, [0.1], [0.0]), -22, "lbm_set_waveform")
        expect(P, lambda: dom.set_waveform(0, P.WAVE_FOURIER, -1.0, [0.1], [0.0]), -22, "lbm_set_waveform")
        expect(P, lambda: dom.macros_box((0, 0, 0), (9, 8, 12)), -22, "bad box")
        expect(P, lambda: dom.macros_box((-1, 0, 0), (2, 2, 2)), -22, "bad box")
        dom.step(5)
        assert dom.iteration == 5
        rho, u = dom.macros()
        assert np.isfinite(rho).all() and np.isfinite(u).all()
        dom.close()


def test_voxelizer_degenerate_inputs(P, ctx):
    from pv2013_amd import cases
    g = cases.lattice_grid((8, 8, 8))
    print("voxelizer degenerate inputs", flush=True)
    # no facets at all: every point is solid
    flags, _ = ctx.voxelize(np.zeros((0, 12)), g)
    assert flags.shape == (512,) and int(flags.sum()) == 0
    # only facets parallel to the +z rays (normal z = 0): none can be crossed
    tri = np.zeros((2, 12))
    tri[0, :9] = [3, 0, 0, 3, 7, 0, 3, 0, 7]
    tri[1, :9] = [3, 7, 7, 3, 7, 0, 3, 0, 7]
    tri[:, 9:] = [1.0, 0.0, 0.0]
    flags, _ = ctx.voxelize(tri, g)
    assert int(flags.sum()) == 0
    # more than 2048 points along z is a range error, and the context stays usable
    big = P.capi.GridDesc()
    for f, _ in P.capi.GridDesc._fields_:
        setattr(big, f, getattr(g, f))
    big.dims[2] = 4096
    expect(P, lambda: ctx.voxelize(tri, big), -34, "nz above 2048")
    flags, _ = ctx.voxelize(tri, g)
    assert int(flags.sum()) == 0


def test_bandwidth_probes(P, ctx):
    """The measured-bandwidth probes: positive rates for valid buffers (the
    lattice-shaped probe in every element size, in place and into a second
    lattice, for a node count that is not a multiple of 64), errors for bad
    arguments."""
    assert ctx.bandwidth(1 << 20, in_place=True, reps=2) > 0
    assert ctx.bandwidth(1 << 20, in_place=False, reps=2) > 0
    for q in (19, 27):
        for eb in (2, 4, 8):
            for pattern in range(5):
                assert ctx.bandwidth_lattice(q, eb, 10007, pattern=pattern, reps=2) > 0
    # a node count one below a multiple of 64: the shifted patterns' last read
    # is the stride's last element
    assert ctx.bandwidth_lattice(19, 4, 4095, pattern=3, reps=1) > 0
    assert ctx.bandwidth_lattice(19, 4, 4095, pattern=4, reps=1) > 0
    expect(P, lambda: ctx.bandwidth_lattice(15, 4, 1000), -22, "bad argument")
    expect(P, lambda: ctx.bandwidth_lattice(19, 3, 1000), -22, "bad argument")
    expect(P, lambda: ctx.bandwidth_lattice(19, 4, 0), -22, "bad argument")
    expect(P, lambda: ctx.bandwidth_lattice(19, 4, 1000, pattern=5), -22, "bad argument")
    assert ctx.bandwidth_lattice(19, 4, 4096, reps=1) > 0
