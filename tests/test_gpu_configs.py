"""The accuracy figures of the benchmark configurations, asserted at the
configurations' own sizes and step counts (FP32 / FP16 shifted storage against
the FP64 GPU path, which the other GPU tests pin to the oracle bit for bit).

Bounds: FP32 velocity L2 <= 1e-5 relative to FP64 (the stated target); FP16
velocity L2 <= 1e-3 and wall shear stress L2 <= 5e-3 (storage rounding of about
2^-11 of the deviations f - w, amplified by the gradient in the WSS)."""
import numpy as np
import pytest

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


def test_c2_pulsatile_8000_steps(P, ctx):
    """Config 2: the 256x256x512 Y-bifurcation with its seeded pulsatile inlet
    (period 4000 steps), 8000 steps in each storage (AA)."""
    from pv2013_amd import cases
    print("building the 256x256x512 Y-bifurcation", flush=True)
    c = cases.c2(ctx)
    fluid = c["flags"] > 0
    wall = c["flags"] == 2
    out = {}
    for name, st in (("f64", P.STORE_F64), ("f32", P.STORE_F32), ("f16s", P.STORE_F16S)):
        print("c2 %s: 8000 steps" % name, flush=True)
        dom = cases.make_domain(ctx, c, storage=st, scheme=P.SCHEME_AA)
        dom.step(8000)
        rho, u = dom.macros()
        out[name] = (rho, u, dom.wss())
        dom.close()
    r64, u64, w64 = out["f64"]
    for name, ub, wb in (("f32", 1e-5, 1e-4), ("f16s", 1e-3, 5e-3)):
        rho, u, w = out[name]
        eu = cases.errors(u64, u, fluid)["l2"]
        ep = cases.errors(r64 / 3.0, rho / 3.0, fluid)["l2"]
        ew = cases.errors(w64, w, wall)["l2"]
        print("c2 %s vs f64: velocity %.3e pressure %.3e wss %.3e" % (name, eu, ep, ew), flush=True)
        assert np.isfinite(u).all() and eu <= ub and ew <= wb and ep <= ub


def test_c5_fp16_wss_against_fp64_at_2000_steps(P, ctx):
    """Config 5: the D3Q27 768x384x384 vessel tree, FP16 shifted storage with
    FP32 collision against FP64, wall shear stress after 2000 steps."""
    from pv2013_amd import cases
    print("building the 768x384x384 D3Q27 vessel tree", flush=True)
    c = cases.c5(ctx)
    wall = c["flags"] == 2
    w = {}
    for name, st in (("f64", P.STORE_F64), ("f16s", P.STORE_F16S)):
        print("c5 %s: 2000 steps" % name, flush=True)
        dom = cases.make_domain(ctx, c, q=27, storage=st, scheme=P.SCHEME_AA)
        dom.step(2000)
        w[name] = dom.wss()
        dom.close()
    e = cases.errors(w["f64"], w["f16s"], wall)
    print("c5 FP16 WSS vs FP64: L2 %.3e Linf %.3e" % (e["l2"], e["linf"]), flush=True)
    assert np.isfinite(w["f16s"]).all() and e["l2"] <= 5e-3


def test_c4_fp32_per_case_error(P, ctx):
    """Config 4: 64 seeded pipes packed into one domain, each with its own
    pulsatile inlet, 4000 steps; every pipe's FP32 velocity within 1e-5 (L2)
    of the FP64 packed run."""
    from pv2013_amd import cases
    print("building the 64-pipe ensemble", flush=True)
    cs = cases.c4(ctx, seed=4, ncases=64)
    us = {}
    for name, st in (("f64", P.STORE_F64), ("f32", P.STORE_F32)):
        dom, info = P.batch(ctx, [dict(dims=c["dims"], flags=c["flags"], tau=c["tau"],
                                       bcs=[P.make_bc(**b) for b in c["bcs"]], q=19, storage=st,
                                       scheme=P.SCHEME_AA) for c in cs])
        for c, inf in zip(cs, info):
            for slot, (kind, period, a, phi) in c["waves"].items():
                dom.set_waveform(inf["wave_base"] + slot, kind, period, a, phi)
        print("c4 %s: %d fluid nodes, 4000 steps" % (name, dom.stats()["nodes"]), flush=True)
        dom.step(4000)
        us[name] = [dom.macros_box(inf["origin"], inf["dims"])[1] for inf in info]
        dom.close()
    errs = [cases.errors(a, b, c["flags"] > 0)["l2"] for a, b, c in zip(us["f64"], us["f32"], cs)]
    print("c4 FP32 per-case velocity L2: max %.3e median %.3e" % (max(errs), float(np.median(errs))), flush=True)
    assert max(errs) <= 1e-5
