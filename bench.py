"""Benchmark of the fused stream-collide step (MLUPS and HBM-roofline fraction).

Workload (one GPU): D3Q19 BGK on a 256x256x512 lattice holding a seeded
Y-bifurcation (parent radius 40 splitting at z=256 into two radius-30
daughters at +-25 degrees, 8-mode wall perturbation of amplitude 1.5, about
150k facets), voxelized on the GPU; Zou-He pulsatile parabolic inlet
u(t) = 0.03 (1 + sum_k a_k sin(2 pi k t / 4000 + phi_k)), Zou-He pressure
outlets rho = 1, half-way bounce-back walls, tau = 0.66.  The timed step
stores populations in FP32 as f - w and collides in FP32.

One JSON line: MLUPS over fluid nodes for the whole job, the roofline of the
step kernel (algorithmic bytes 2*Q*s per fluid node per step over the average
launch time measured with HIP events on the engine's stream), and the CPU
oracle timed on a bounded sample of the same workload on this host.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--storage f32|f64|f16s]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
INFINITY_CACHE_BYTES = 256 * 2 ** 20  # MI355X last-level (MALL) cache
TAU = 0.66
SEED = 2013
PERIOD = 4000.0


def run(P, ctx, case, storage, steps, warmup, scheme=0, options=0, block=0):
    from pv2013_amd.cases import make_domain
    dom = make_domain(ctx, case, storage=storage, scheme=scheme, options=options, block=block)
    st = dom.stats()
    n = st["nodes"]
    dom.step(warmup)
    dom.sync()
    t0 = time.perf_counter()
    ms = dom.step_timed(steps)
    wall = time.perf_counter() - t0
    mass, _ = dom.reduce()
    dom.close()
    return {"nodes": n, "ms": ms, "wall_s": wall, "mass": mass, "stats": st}


def accuracy(P, ctx, case, storages):
    """The configuration's protocol: 8000 steps from rest in each storage;
    velocity, pressure (rho/3) and wall-shear-stress errors against FP64."""
    from pv2013_amd.cases import errors, make_domain
    out = {}
    ref = None
    for name, st in (("f64", P.STORE_F64),) + tuple(storages):
        dom = make_domain(ctx, case, storage=st)
        dom.step(case["steps"])
        rho, u = dom.macros()
        wss = dom.wss()
        dom.close()
        if ref is None:
            ref = (rho, u, wss)
            fl = case["flags"] > 0
            wl = case["flags"] == 2
            continue
        out[name] = {"velocity": errors(ref[1], u, fl), "pressure": errors(ref[0] / 3.0, rho / 3.0, fl),
                     "wss": errors(ref[2], wss, wl), "steps": case["steps"]}
    return out


def cpu_baseline(case, budget_s=10.0, max_threads=16):
    """The FP64 oracle (a plain-C restatement of the reference step) on the
    same domain, for as many steps as fit in about budget_s: once on one
    thread, once with OpenMP on the host cores given to this GPU (at most 16).
    The multi-core figure is the baseline; the 1-thread one goes in the sample
    text."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    bcs = [O.make_bc(**b) for b in case["bcs"]]
    sim = O.Sim(19, case["dims"], case["flags"], TAU, bcs, quirks=1)
    kind, period, a, phi = case["waves"][0]
    sim.set_waveform(0, kind, period, a, phi)
    threads = max(1, min(max_threads, len(os.sched_getaffinity(0)), O.threads_available()))

    def timed(step):
        t0 = time.perf_counter()
        step(1)
        one = time.perf_counter() - t0
        k = max(1, int(budget_s / max(one, 1e-6)) - 1)
        t0 = time.perf_counter()
        step(k)
        return k, sim.nodes * k / (time.perf_counter() - t0) / 1e6

    k1, one_thread = timed(sim.step)
    kn, many = timed(lambda k: sim.step_threads(k, threads))
    return {"value": many, "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": "%d steps of the same 256x256x512 Y-bifurcation domain (%d fluid nodes), FP64 oracle, "
                      "OpenMP on %d threads; %d steps on 1 thread: %.2f MLUPS" % (kn, sim.nodes, threads, k1, one_thread)}


def traffic_from_profiles(key):
    """HBM bytes per step measured with rocprofv3 FETCH_SIZE / WRITE_SIZE for
    this storage / scheme / layout, and where that figure comes from
    (profiles/step_traffic.json: the profile file, command and commit), or
    (None, None).  Counters cannot be read from inside the run, so the figure
    is the profiled run's, not this run's."""
    path = os.path.join(ROOT, "profiles", "step_traffic.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None, None
    if key not in t:
        return None, None
    return t[key], t.get("_source")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    # At rest the stored deviations f - w are exact zeros, and a lattice of
    # zeros streams about 15% faster than one holding a flow; the inlet's
    # pressure wave needs about 900 steps to cross the 512-node vessel, so the
    # default warm-up lets the flow fill the lattice before the clock starts.
    ap.add_argument("--warmup", type=int, default=1500)
    ap.add_argument("--storage", default="f32", choices=["f32", "f64", "f16s"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the FP64 / FP16 comparison runs and the accuracy protocol")
    ap.add_argument("--scheme", default="aa", choices=["two", "aa"],
                    help="streaming: AA (one lattice, default) or two-lattice pull; identical results")
    ap.add_argument("--block", type=int, default=0, choices=[0, 64, 128, 256, 512],
                    help="threads per block of the step kernels (0: the library default: 64 for FP16 "
                         "storage; two-lattice 128 for FP32, 256 for FP64; AA 512 while the lattice fits "
                         "the Infinity Cache, else 256)")
    ap.add_argument("--layout", default="fused", choices=["fused", "split", "generic"],
                    help="step kernel: row chunks in one launch (default), records + row segments "
                         "in two, or the generic per-node kernel")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the whole domain runs on a single MI355X "
              "(multi-GPU sharding is a design note, not built); run with --gpus 1", file=sys.stderr)
        return 2
    import pv2013_amd as P
    storage = {"f32": P.STORE_F32, "f64": P.STORE_F64, "f16s": P.STORE_F16S}[args.storage]
    sbytes = {"f32": 4, "f64": 8, "f16s": 2}[args.storage]
    ctx = P.Context(0)
    from pv2013_amd.cases import c2
    print("bench: building the 256x256x512 Y-bifurcation case", file=sys.stderr, flush=True)
    case = c2(ctx, seed=SEED, tau=TAU, period=PERIOD)
    case["nsites"] = 256 * 256 * 512
    print("bench: timing %d steps (%s) after %d warm-up" % (args.steps, args.storage, args.warmup),
          file=sys.stderr, flush=True)
    scheme = P.SCHEME_AA if args.scheme == "aa" else P.SCHEME_TWO_LATTICE
    options = {"fused": 0, "split": P.capi.OPT_SPLIT_LAUNCH, "generic": P.capi.OPT_GENERIC_KERNEL}[args.layout]
    r = run(P, ctx, case, storage, args.steps, args.warmup, scheme=scheme, options=options, block=args.block)
    n = r["nodes"]
    ms_step = r["ms"] / args.steps
    mlups = n * args.steps / (r["ms"] * 1e-3) / 1e6
    bytes_per_launch = n * 2 * 19 * sbytes
    achieved = bytes_per_launch / (ms_step * 1e-3) / 1e9
    extra = {"fluid_nodes": n, "box_sites": case["nsites"], "box_site_mlups": case["nsites"] / n * mlups,
             "wall_s_timed": r["wall_s"], "mass_after": r["mass"], "facets": case["ntri"],
             "voxelize_ms": case["vox_ms"], "classify_ms": case["cls_ms"], "scheme": args.scheme,
             "voxels_per_s": case["nsites"] / (case["vox_ms"] * 1e-3),
             "layout": args.layout, "block": r["stats"]["block"], "boundary_records": r["stats"]["boundary_records"],
             "row_chunks" if args.layout == "fused" else "segments": r["stats"]["segments"],
             "step_waves": r["stats"]["step_waves"],
             "lane_use": ((n if args.layout == "fused" else n - r["stats"]["boundary_records"])
                          / max(1, 64 * r["stats"]["step_waves"])),
             "node_order_fast_axis": r["stats"]["fast_axis"]}
    if args.layout == "generic":
        extra["regular_wave_fraction"] = r["stats"]["regular_waves"] / max(1, r["stats"]["waves"])
    # measured device bandwidth next to the published peak: in place over the
    # size of the lattice (the AA scheme's access pattern), a 1 GiB copy, and
    # the step's own arrays moved with aligned accesses and no arithmetic (the
    # practical ceiling of this layout), in place and into a second lattice
    probes = {k: ctx.bandwidth_lattice(19, sbytes, n, pattern=v, reps=10)
              for k, v in (("in_place", P.capi.PROBE_IN_PLACE), ("in_place_shifted", P.capi.PROBE_IN_PLACE_SHIFTED),
                           ("two", P.capi.PROBE_TWO), ("two_shifted", P.capi.PROBE_TWO_SHIFTED),
                           ("aa", P.capi.PROBE_AA))}
    # the AA step alternates the gather (the shifted in-place pattern, walking
    # downward) and the local step (in place, aligned, walking upward)
    probe_aa = probes["aa"]
    probe_two = probes["two_shifted"]
    extra["measured_gbs"] = {"in_place_lattice_size": ctx.bandwidth(n * 19 * sbytes, in_place=True),
                             "copy_1GiB": ctx.bandwidth(1 << 30, in_place=False, reps=5),
                             "lattice_probes": probes}
    # The AA lattice of this case (19 x 4 B x 2.74M nodes = 209 MB) fits in the
    # 256 MB Infinity Cache, so the headline step is not purely HBM-streaming.
    # The same storage with two lattices (418 MB, larger than the cache) runs
    # beside it as the HBM-streaming figure.
    lattice_bytes = 19 * sbytes * r["stats"]["nodes"] * (1 if scheme == P.SCHEME_AA else 2)
    streaming = None
    if scheme == P.SCHEME_AA and not args.no_compare:
        print("bench: HBM-streaming run (two lattices)", file=sys.stderr, flush=True)
        rs = run(P, ctx, case, storage, min(args.steps, 1000), args.warmup, scheme=P.SCHEME_TWO_LATTICE,
                 options=options, block=args.block)
        ms2 = rs["ms"] / min(args.steps, 1000)
        ach2 = bytes_per_launch / (ms2 * 1e-3) / 1e9
        streaming = {"scheme": "two-lattice pull", "lattice_bytes": 2 * 19 * sbytes * rs["nodes"],
                     "cache_resident": False, "mlups": rs["nodes"] / (ms2 * 1e-3) / 1e6,
                     "achieved": ach2, "frac": ach2 / HBM_PEAK_GBS,
                     "practical_ceiling": {"gbs": probe_two, "frac": ach2 / probe_two,
                                           "aligned_gbs": probes["two"], "aligned_frac": ach2 / probes["two"]}}
    if not args.no_compare:
        for name, st in (("f64", P.STORE_F64), ("f16s", P.STORE_F16S)):
            if name == args.storage:
                continue
            # each storage in both streaming schemes; its figure is the faster
            # (the FP64 pull is faster than FP64 AA on this case, so the
            # speed-up is taken against the best FP64 path, not the AA one)
            by = {}
            for sname, sch in (("aa", P.SCHEME_AA), ("two", P.SCHEME_TWO_LATTICE)):
                print("bench: comparison run %s %s" % (name, sname), file=sys.stderr, flush=True)
                rr = run(P, ctx, case, st, min(args.steps, 1000), args.warmup, scheme=sch, options=options,
                         block=args.block)
                by[sname] = rr["nodes"] * min(args.steps, 1000) / (rr["ms"] * 1e-3) / 1e6
            extra["mlups_" + name] = max(by.values())
            extra["mlups_%s_by_scheme" % name] = by
        if "mlups_f64" in extra:
            extra["speedup_vs_f64"] = mlups / extra["mlups_f64"]
        print("bench: accuracy protocol (%d steps per storage)" % case["steps"], file=sys.stderr, flush=True)
        extra["errors_vs_f64"] = accuracy(P, ctx, case, (("f32", P.STORE_F32), ("f16s", P.STORE_F16S)))
    traffic, traffic_src = traffic_from_profiles("%s_%s_%s" % (args.storage, args.scheme, args.layout))
    cpu = None
    if not args.no_cpu:
        print("bench: CPU oracle sample", file=sys.stderr, flush=True)
        cpu = cpu_baseline(case)
    line = {
        "metric": "MLUPS & % HBM roofline, D3Q19 256^2x512; FP32/FP16 vs FP64 at stated error",
        "value": mlups,
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.storage == "f64" else "f32",
        "data": "synthetic (seeded Y-bifurcation STL, seed %d)" % SEED,
        "config": {"workload": "D3Q19 BGK, 256x256x512 Y-bifurcation, pulsatile Zou-He inlet, "
                               "pressure outlets, bounce-back walls, tau 0.66",
                   "lattice": "D3Q19", "box": [256, 256, 512],
                   "storage": {"f32": "f32 holding f - w", "f64": "f64",
                               "f16s": "f16 holding f - w"}[args.storage],
                   "scheme": "AA" if args.scheme == "aa" else "two-lattice pull",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "lattice_bytes": lattice_bytes,
                     "cache_resident": lattice_bytes <= INFINITY_CACHE_BYTES,
                     "practical_ceiling": {
                         "probe": "lbm_ctx_bandwidth_lattice: the step's 19 arrays of this node count and storage "
                                  "read then written with no arithmetic, "
                                  + ("in place, alternating a pass with every run one node off its line walking "
                                     "the lattice downward (the AA gather's pattern and order) and an aligned pass "
                                     "walking it upward (the local step's)" if scheme == P.SCHEME_AA
                                     else "into a second lattice, reads one node off their lines (the pull's "
                                          "pattern)") + "; aligned_*: every access aligned",
                         "gbs": probe_aa if scheme == P.SCHEME_AA else probe_two,
                         "frac": achieved / (probe_aa if scheme == P.SCHEME_AA else probe_two),
                         "aligned_gbs": probes["in_place"] if scheme == P.SCHEME_AA else probes["two"],
                         "aligned_frac": achieved / (probes["in_place"] if scheme == P.SCHEME_AA else probes["two"])},
                     "hbm_streaming": streaming},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line))
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
