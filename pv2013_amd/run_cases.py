"""Runs the non-headline benchmark configurations on the GPU and prints one
JSON line per case (the headline 256x256x512 case is bench.py).

  python -m pv2013_amd.run_cases c3 c4 c5 [--quick]

c3  voxelizer only: seeded tortuous tube (~540k facets) into a 512^3 grid;
    voxels/s of the device work (facet binning and voxelizer kernels) and of
    the whole lbm_voxelize call (facet upload and 134 MB flag download
    included), and whether repeated runs give the same flags (the GPU tests
    check this mesh's flags against the CPU oracle).
c4  64 seeded straight and bent pipes (lengths 64-512, radii 8-40, curvature
    radius 100-1000), FP32 storage, each with its own pulsatile inlet, packed
    into one domain (one launch per step, hipGraph chunks), 4000 steps:
    aggregate MLUPS and per-case velocity error against the same cases in FP64.
c5  D3Q27 768x384x384 vessel tree (3 bifurcation levels, radii 60 -> 15), FP16
    shifted storage with FP32 collision, 10000 steps: MLUPS (AA, and the
    two-lattice pull timed on the same developed flow); wall shear stress
    error against an FP64 run, both after 2000 steps.
c4 and c5 also report the practical ceiling: the step's own arrays moved
in the step's access pattern with no arithmetic (lbm_ctx_bandwidth_lattice).
"""
import argparse
import json
import sys
import time

import numpy as np

from . import capi as P
from . import cases

HBM = 8000.0


def run_c3(ctx, quick):
    n = 256 if quick else 512
    t0 = time.perf_counter()
    tri, g = cases.c3_mesh(seed=3, n=n)
    gen = time.perf_counter() - t0
    flags, ms = ctx.voxelize(tri, g)
    calls = []
    for _ in range(3):  # warm calls: the whole lbm_voxelize, host copies included
        t0 = time.perf_counter()
        flags2, ms2 = ctx.voxelize(tri, g)
        calls.append((time.perf_counter() - t0, ms2))
        if not np.array_equal(flags, flags2):
            break
    call_s, dev_ms = min(calls)
    out = {"case": "c3", "grid": [n, n, n], "facets": len(tri), "generate_s": gen,
           "device_ms": dev_ms, "voxels_per_s_device": n ** 3 / (dev_ms * 1e-3),
           "call_ms": call_s * 1e3, "voxels_per_s_call": n ** 3 / call_s,
           "fluid_fraction": float(flags.mean()), "repeat_identical": bool(np.array_equal(flags, flags2))}
    return out


def run_c4(ctx, quick):
    ncases = 8 if quick else 64
    steps = 400 if quick else 4000
    t0 = time.perf_counter()
    cs = cases.c4(ctx, seed=4, ncases=ncases)
    gen = time.perf_counter() - t0
    res = {}
    for name, st in (("f32", P.STORE_F32), ("f64", P.STORE_F64)):
        dom, info = P.batch(ctx, [dict(dims=c["dims"], flags=c["flags"], tau=c["tau"],
                                       bcs=[P.make_bc(**b) for b in c["bcs"]], q=19, storage=st,
                                       scheme=P.SCHEME_AA)
                                  for c in cs])
        for c, inf in zip(cs, info):
            for slot, (kind, period, a, phi) in c["waves"].items():
                dom.set_waveform(inf["wave_base"] + slot, kind, period, a, phi)
        stats = dom.stats()
        dom.step(64)
        ms = dom.step_timed(steps - 64)
        per_case = [dom.macros_box(inf["origin"], inf["dims"])[1] for inf in info]
        res[name] = {"mlups": stats["nodes"] * (steps - 64) / (ms * 1e-3) / 1e6, "nodes": stats["nodes"],
                     "u": per_case,
                     "lane_use": stats["nodes"] / max(1, 64 * stats["step_waves"])}
        dom.close()
    errs = [cases.errors(u64, u32, c["flags"] > 0)["l2"]
            for u64, u32, c in zip(res["f64"]["u"], res["f32"]["u"], cs)]
    # the packed FP64 run equals each case run alone, bit for bit: check one
    c = cs[0]
    single = cases.make_domain(ctx, c, storage=P.STORE_F64, scheme=P.SCHEME_AA)
    single.step(steps)
    same = bool(np.array_equal(single.macros()[1], res["f64"]["u"][0]))
    single.close()
    # the step's own arrays moved with no arithmetic, the AA pattern (gather
    # runs one node off their lines, local step aligned): the practical
    # ceiling of this layout on this node count
    ceil = aa_ceiling(ctx, 19, 4, res["f32"]["nodes"])
    return {"case": "c4", "cases": ncases, "steps": steps, "generate_s": gen, "fluid_nodes": res["f32"]["nodes"],
            "mlups_f32": res["f32"]["mlups"], "mlups_f64": res["f64"]["mlups"],
            "roofline_frac_f32": res["f32"]["mlups"] * 152 / 1e3 / HBM,
            "practical_ceiling_gbs": ceil, "practical_ceiling_frac_f32": res["f32"]["mlups"] * 152 / 1e3 / ceil,
            "velocity_l2_vs_f64": {"max": max(errs), "median": float(np.median(errs))},
            "packed_f64_equals_single_case": same,
            "lane_use": res["f32"]["lane_use"]}


def aa_ceiling(ctx, q, eb, nodes):
    """GB/s of the AA step's traffic with no arithmetic: pairs of in-place
    passes, a shifted one walking downward (the gather's runs start one node
    off their lines) and an aligned one walking upward (the local step)."""
    return ctx.bandwidth_lattice(q, eb, nodes, pattern=P.PROBE_AA, reps=5)


def run_c5(ctx, quick, steps=None):
    dims = (384, 192, 192) if quick else (768, 384, 384)
    steps = steps or (1000 if quick else 10000)
    cmp_steps = min(500 if quick else 2000, steps // 2)
    t0 = time.perf_counter()
    c = cases.c5(ctx, seed=5, dims=dims, h=4.0 if quick else 3.0)
    gen = time.perf_counter() - t0
    dom = cases.make_domain(ctx, c, q=27, storage=P.STORE_F16S, scheme=P.SCHEME_AA)
    st = dom.stats()
    dom.step(cmp_steps)
    w16 = dom.wss()
    ms = dom.step_timed(steps - cmp_steps)
    mass, _ = dom.reduce()
    dom.close()
    # the two-lattice pull (chunk pairs) on the same case, in the same
    # developed-flow regime: warmed up for cmp_steps, then timed
    two = cases.make_domain(ctx, c, q=27, storage=P.STORE_F16S, scheme=P.SCHEME_TWO_LATTICE)
    st2 = two.stats()
    two.step(cmp_steps)
    n2 = min(steps - cmp_steps, 2000)
    ms2 = two.step_timed(n2)
    two.close()
    ref = cases.make_domain(ctx, c, q=27, storage=P.STORE_F64, scheme=P.SCHEME_AA)
    ref.step(cmp_steps)
    w64 = ref.wss()
    ms64 = ref.step_timed(200)
    ref.close()
    wall = c["flags"] == 2
    mlups = st["nodes"] * (steps - cmp_steps) / (ms * 1e-3) / 1e6
    mlups2 = st2["nodes"] * n2 / (ms2 * 1e-3) / 1e6
    # practical ceilings: the step's 27 FP16 arrays moved with no arithmetic,
    # in the AA pattern and the pull's (reads one node off their lines)
    ceil_aa = aa_ceiling(ctx, 27, 2, st["nodes"])
    ceil_two = ctx.bandwidth_lattice(27, 2, st["nodes"], pattern=P.PROBE_TWO_SHIFTED, reps=5)
    return {"case": "c5", "box": list(dims), "facets": c["ntri"], "generate_s": gen, "fluid_nodes": st["nodes"],
            "steps": steps, "mlups_f16s": mlups, "roofline_frac_f16s": mlups * 108 / 1e3 / HBM,
            "practical_ceiling_gbs": ceil_aa, "practical_ceiling_frac_f16s": mlups * 108 / 1e3 / ceil_aa,
            "two_lattice": {"mlups_f16s": mlups2, "timed_steps": n2, "step_waves": st2["step_waves"],
                            "lane_use": st2["nodes"] / (64.0 * st2["step_waves"]),
                            "roofline_frac_f16s": mlups2 * 108 / 1e3 / HBM, "practical_ceiling_gbs": ceil_two,
                            "practical_ceiling_frac_f16s": mlups2 * 108 / 1e3 / ceil_two},
            "mlups_f64": st["nodes"] * 200 / (ms64 * 1e-3) / 1e6,
            "wss_error_vs_f64_at_%d" % cmp_steps: cases.errors(w64, w16, wall), "mass_after": mass}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cases", nargs="+", choices=["c3", "c4", "c5"])
    ap.add_argument("--quick", action="store_true", help="reduced sizes for a smoke run")
    ap.add_argument("--steps", type=int, default=None,
                    help="c5 only: total FP16 steps (default 10000; for short profiled runs)")
    args = ap.parse_args()
    ctx = P.Context(0)
    for name in args.cases:
        print("run_cases: %s" % name, file=sys.stderr, flush=True)
        if name == "c5":
            res = run_c5(ctx, args.quick, args.steps)
        else:
            res = {"c3": run_c3, "c4": run_c4}[name](ctx, args.quick)
        print(json.dumps(res, default=float), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
