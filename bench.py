"""Benchmark: options/s of the FP64 trinomial-tree pricer on one MI355X.

Workload: the 60,000-option skewed batch (90% small trees: bond 1-5y at
12 steps/yr, a in [0.1,0.2]; 10% large trees: bond 10-30y at 36 steps/yr,
a in [0.01,0.05], widths up to 1,327 nodes, heights up to 1,080 steps),
synthetic 120-knot yield curve, seeded inputs already resident in HBM.
A step prices the whole batch: device-side constants and ordering plus the two
pricing kernels, replayed from the plan's HIP graph or launched directly
(HWT_FLAG_DIRECT), whichever the autotuner measured faster.

Prints one JSON line.  `roofline` is for the dominant kernel (the wide-tree
pricing kernel, timed alone with HIP events on its stream): FP64 vector FLOPs
of the stencil (5 per node update: one multiply and two fused multiply-adds)
against MI355X's 78.6 TFLOP/s FP64 peak.  `cpu_baseline` times the reference's
own sequential pricer (built from its sources into oracle/_ref) on a bounded
stratified sample on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "options/sec on 1 MI355X: FP64 tree vs surrogate+fallback at max err 1e-3"
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
FLOP_PER_NODE_UPDATE = 5


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(pkg, opts, y, c1):
    """The reference's seq binary (REAL=64) on the same 60,000-option batch,
    timed by its own -r mode (best of the runs, seq::computeOptions only, no
    file I/O): with one OpenMP thread per host core (one option per core at a
    time, best of three) and single-threaded (one run, about 20 s of CPU work
    on a current server core); and single-threaded on the 1,000-put workload
    (best of three).  Without the reference binary the oracle's C restatement
    of seq is timed instead ("port")."""
    import oracle  # the reference binary lives beside the oracle (oracle/_ref)
    seq = oracle.REF_SEQ
    kind = "reference" if seq.exists() else "port"
    ncores = os.cpu_count() or 1
    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        pkg.formats.write_yield(td / "yield.in", y)

        def run(sample, threads, runs):
            pkg.formats.write_options(td / "opts.in", sample)
            env = {**os.environ, "OMP_NUM_THREADS": str(threads)}
            if seq.exists():
                out = subprocess.run([str(seq), "-o", str(td / "opts.in"), "-y", str(td / "yield.in"),
                                      "-r", str(runs)], check=True, capture_output=True, text=True, env=env).stdout
                return int(out.strip().split(",")[4]) * 1e-6  # "-,-,-,-,<best us>,-"
            best = None
            for _ in range(runs):
                t0 = time.perf_counter()
                oracle.price(sample, y, threads=threads)
                t = time.perf_counter() - t0
                best = t if best is None else min(best, t)
            return best

        t_multi = run(opts, ncores, 3)
        t_single = run(opts, 1, 1)
        t_c1 = run(c1, 1, 3)
    what = "reference seq (REAL=64, OpenMP)" if kind == "reference" else "oracle restatement of seq"
    return {"value": opts.N / t_multi, "unit": "options/s", "cores": ncores, "kind": kind,
            "cpu_model": _cpu_model(),
            "sample": f"the whole 60k skewed batch, {what} with {ncores} threads, best of 3 runs: {t_multi:.3f} s "
                      f"(seq::computeOptions only, timed by the binary's -r mode)",
            "single_thread_value": opts.N / t_single,
            "single_thread_sample": f"the whole 60k skewed batch, 1 thread, one run: {t_single:.2f} s",
            "config1_single_thread_value": c1.N / t_c1,
            "config1_sample": f"1,000 puts (width 47, 120 steps), 1 thread, best of 3: {t_c1 * 1e3:.1f} ms"}


def reference_sets_leg(pkg, capi, args):
    """The reference's seven 65,536-option benchmark distributions (drawn by
    the seeded generator; the reference's 15-knot curve), each priced whole
    by the FP64 tree: back-to-back executions timed with HIP events, in the
    autotuned configuration (graph replays or direct launches).  The reference's own figures for these sets are on a
    GTX 780 Ti (CUDA-multi FP64, 4_SKEWED: 1.147 s)."""
    import ctypes as C
    dm = pkg.devmem
    yref = pkg.formats.read_yield(ROOT / "tests" / "golden" / "yield.in")
    stream = C.c_void_p()
    dm._ck(dm.hip().hipStreamCreate(C.byref(stream)))
    out = {}
    for name in pkg.datasets.REFERENCE_SETS:
        opts = pkg.datasets.reference_set(name, 65536, seed=5)
        dev = capi.to_device(opts, yref)
        cfg, _ = capi.autotune(dev)
        plan = capi.Plan(opts.N, yref.N, cfg)
        for _ in range(3):
            plan.execute(dev, stream=stream.value)
        st = plan.sync()
        e0, e1 = dm.Event(), dm.Event()
        reps = 10
        e0.record(stream)
        for _ in range(reps):
            plan.execute(dev, stream=stream.value)
        e1.record(stream)
        e1.synchronize()
        plan.sync()
        plan.close()
        ms = e0.elapsed_ms(e1) / reps
        out[name] = {"ms": ms, "options_per_s": opts.N / (ms * 1e-3), "node_updates": st["node_updates"],
                     "seq_node_updates": st["seq_node_updates"], "tuned_form": cfg.form, "tuned_flags": cfg.flags}
        # form 1 (one lane per option, tree columns in HBM) on the same batch,
        # pricing kernel alone: 16 B of column traffic per node update (the
        # next value read, the new value written)
        plan = capi.Plan(opts.N, yref.N, capi.config(form=capi.FORM_LANE, flags=capi.FLAG_NO_GRAPH))
        plan.execute(dev)
        s1 = plan.sync()
        kms = []
        for _ in range(3 if s1["kernel_ms"] < 50 else 0):
            plan.execute(dev)
            kms.append(plan.sync()["kernel_ms"])
        plan.close()
        if kms:
            km = float(np.median(kms))
            gbs = 16.0 * s1["node_updates"] / (km * 1e-3) / 1e9
            out[name]["form1"] = {"kernel_ms": km, "options_per_s": opts.N / (km * 1e-3),
                                  "column_gb_per_s": gbs, "hbm_frac": gbs / HBM_PEAK_GBS}
        else:
            out[name]["form1"] = {"kernel_ms_first": s1["kernel_ms"], "note": "not repeated (over 50 ms)"}
        for k in list(dev):
            if hasattr(dev[k], "free"):
                dev[k].free()
    dm._ck(dm.hip().hipStreamDestroy(stream))
    return out


def tiny_trees_leg(pkg, capi, y, n: int = 1_000_000):
    """The flattened form against the row form on a batch of tiny trees
    (widths 3-21, uniform): the same options priced with trees packed 64 / w
    to a wavefront one node per lane (default), four per wavefront in 16-lane
    rows (HWT_FLAG_NO_FLAT), and one per wavefront (also HWT_FLAG_NO_ROWS).
    Graph replays timed with HIP events."""
    import ctypes as C
    dm = pkg.devmem
    opts = pkg.datasets.tiny(7, n)
    dev = capi.to_device(opts, y)
    stream = C.c_void_p()
    dm._ck(dm.hip().hipStreamCreate(C.byref(stream)))
    out = {"options": n}
    for name, flags in (("flat", 0), ("rows", capi.FLAG_NO_FLAT),
                        ("one_per_wave", capi.FLAG_NO_FLAT | capi.FLAG_NO_ROWS)):
        plan = capi.Plan(n, y.N, capi.config(flags=flags))
        for _ in range(3):
            plan.execute(dev, stream=stream.value)
        st = plan.sync()
        e0, e1 = dm.Event(), dm.Event()
        e0.record(stream)
        for _ in range(10):
            plan.execute(dev, stream=stream.value)
        e1.record(stream)
        e1.synchronize()
        plan.sync()
        plan.close()
        ms = e0.elapsed_ms(e1) / 10
        out[name] = {"ms": ms, "options_per_s": n / (ms * 1e-3), "tasks": st["tasks"]}
    out["flat_speedup_over_rows"] = out["rows"]["ms"] / out["flat"]["ms"]
    dm._ck(dm.hip().hipStreamDestroy(stream))
    for k in list(dev):
        if hasattr(dev[k], "free"):
            dev[k].free()
    return out


def surrogate_leg(pkg, y, args):
    """Surrogate + fallback against the plain FP64 tree on the 1M-option mix
    (95% inside the training box, 5% outside).  The ensemble (4 members,
    8 -> 256 x 4 -> 1) is first trained here, on the GPU, on 2M box samples
    priced by the FP64 tree; training is reported, not timed as a step."""
    S = pkg.surrogate
    s, fit = S.fit(y, n_train=2_000_000, epochs=args.surrogate_epochs, seed=11,
                   log=lambda m: print(m, file=sys.stderr, flush=True))
    # the gate's spread threshold, calibrated on a held-out 1M sample (its own
    # seed) so that surrogate-priced options stay within half the 1e-3 budget
    # of the tree.  Calibrated on 250,000 options at 0.7e-3, one network let
    # an option through at 1.46e-3 on the evaluation mix; at 1M and 0.5e-3
    # the same network's worst accepted option was 0.49e-3.
    cal = S.calibrate_threshold(s, y, n=1_000_000, seed=77, target=1e-3, margin=0.5)
    gated = S.with_threshold(s, cal["threshold"], 1 << 20, y.N)
    ev = S.evaluate(gated, y, n=1_000_000, seed=5, reps=max(1, min(args.steps, 10)))
    gated.close()
    s.close()
    # the inference kernel evaluates the in-box options only (the box split
    # sends the others straight to the tree)
    in_box = ev["options"] - ev["out_of_box"]
    flop = in_box * 4 * 2 * (8 * 256 + 3 * 256 * 256 + 256)
    achieved = flop / (ev["inference_ms"] * 1e-3) / 1e12
    return {"workload": "1,000,000 options: 95% inside the training box, 5% outside (a in [0.2,0.4] "
                        "or expiry 30-40y); ensemble BF16 MLP inference + FP64 tree fallback",
            "training": fit, "calibration": cal, **ev,
            "roofline": {"bound": "mfma", "kernel": "k_infer", "achieved": achieved, "peak": 2500.0,
                         "unit": "TFLOP/s", "frac": achieved / 2500.0, "traffic": None,
                         "flop_per_launch": flop,
                         "flop_model": "4 members x 2 x (8*256 + 3*256*256 + 256) FLOP per in-box option "
                                       "(the rows the kernel evaluates); time = box split + inference; "
                                       "BF16 MFMA dense peak"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-reference-sets", action="store_true",
                    help="skip the seven 65,536-option reference distributions in the detail")
    ap.add_argument("--surrogate-epochs", type=int, default=120,
                    help="training epochs of the surrogate ensemble over its 2M tree-priced samples")
    ap.add_argument("--profile", action="store_true",
                    help="profiler run: one 64 MiB device-to-device copy (the byte-counter calibration "
                         "dispatch), then warmup + steps executions of the pipeline without the graph, "
                         "with the settings given by --sort/--concurrent/--no-rows (no tuning runs, so every "
                         "traced pricing dispatch is of the profiled configuration)")
    ap.add_argument("--sort", default="L", help="--profile: processing order of the wide trees")
    ap.add_argument("--concurrent", action="store_true", help="--profile: narrow-tree kernel beside the wide one")
    ap.add_argument("--no-rows", action="store_true", help="--profile: narrow trees one per wavefront")
    ap.add_argument("--fused", action="store_true", help="--profile: one persistent kernel for every tree")
    ap.add_argument("--mixed", action="store_true", help="--profile: mixed precision (FP32 tree passes)")
    ap.add_argument("--profile-lane", default=None, metavar="SET",
                    help="form 1 alone, without the graph, on one of the reference's 65,536-option sets "
                         "(e.g. 0_UNIFORM), for a profiler")
    ap.add_argument("--profile-surrogate", action="store_true",
                    help="profiler run of the surrogate: a short training (2 epochs on 262,144 tree-priced "
                         "samples), then warmup + steps inference-with-fallback passes over the 1M-option mix")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): options are independent, "
              "so a multi-GPU run is a disjoint split of the batch; see the sharding plan in DESIGN.md",
              file=sys.stderr)
        sys.exit(2)

    import __graft_entry__ as ge
    pkg = ge.load_package()
    capi, dm = pkg.capi, pkg.devmem
    y = pkg.datasets.synthetic_yield()
    if args.profile_lane:
        yref = pkg.formats.read_yield(ROOT / "tests" / "golden" / "yield.in")
        opts = pkg.datasets.reference_set(args.profile_lane, 65536, seed=5)
        dev = capi.to_device(opts, yref)
        plan = capi.Plan(opts.N, yref.N, capi.config(form=capi.FORM_LANE, flags=capi.FLAG_NO_GRAPH))
        ks = []
        for i in range(args.warmup + args.steps):
            plan.execute(dev)
            st = plan.sync()
            if i >= args.warmup:
                ks.append(st["kernel_ms"])
            print(f"profile lane step {i}: form-1 kernel {st['kernel_ms']:.4f} ms", flush=True)
        plan.close()
        km = float(np.mean(ks))
        print(json.dumps({"profile_lane": args.profile_lane, "options": opts.N, "kernel_ms_mean": km,
                          "node_updates": st["node_updates"],
                          "column_gb_per_s": 16.0 * st["node_updates"] / (km * 1e-3) / 1e9}), flush=True)
        return
    if args.profile_surrogate:
        S = pkg.surrogate
        s, _ = S.fit(y, n_train=262144, epochs=2, seed=11)
        mix = pkg.datasets.box_mix(5, 1_000_000)
        dmix = capi.to_device(mix, y)
        fb = capi.Plan(mix.N, y.N, capi.config(flags=capi.FLAG_NO_GRAPH))
        for i in range(args.warmup + args.steps):
            st = s.price(fb, dmix)
            print(f"profile surrogate pass {i}: {st['ms']:.3f} ms (inference {st['surrogate_ms']:.3f} ms, "
                  f"fallback {st['fallback']})", flush=True)
        fb.close()
        s.close()
        return
    opts = pkg.datasets.skewed_60k()
    dev = capi.to_device(opts, y)

    if args.profile:
        n = 64 << 20
        src, dst = dm.DeviceArray(n, np.uint8, (n,)), dm.DeviceArray(n, np.uint8, (n,))
        dm._ck(dm.hip().hipMemcpy(dst.ptr, src.ptr, n, dm.D2D))
        dm.synchronize()
        flags = capi.FLAG_NO_GRAPH | (capi.FLAG_CONCURRENT if args.concurrent else 0) | \
            (capi.FLAG_NO_ROWS if args.no_rows else 0) | (capi.FLAG_FUSED if args.fused else 0)
        prec = capi.PREC_MIXED if args.mixed else capi.PREC_FP64
        plan = capi.Plan(opts.N, y.N, capi.config(sort=args.sort, flags=flags, precision=prec))
        wide = []
        for i in range(args.warmup + args.steps):
            plan.execute(dev)
            st = plan.sync()
            if i >= args.warmup:
                wide.append(st["wide_ms"])
            print(f"profile step {i}: wide kernel {st['wide_ms']:.4f} ms", flush=True)
        plan.close()
        print(json.dumps({"profile": "skewed_60k", "precision": "mixed" if args.mixed else "fp64", "copy_bytes": n,
                          "wide_ms_mean": float(np.mean(wide)), "wide_node_updates": st["wide_node_updates"],
                          "fp64_repriced": st["fp64_repriced"]}), flush=True)
        return

    # whole-job throughput: executions on a caller stream (graph replays or direct
    # launches, as tuned), HIP events around K steps
    import ctypes as C
    stream = C.c_void_p()
    dm._ck(dm.hip().hipStreamCreate(C.byref(stream)))

    def time_steps(cfg):
        plan = capi.Plan(opts.N, y.N, cfg)
        for _ in range(args.warmup):
            plan.execute(dev, stream=stream.value)
        plan.sync()
        e0, e1 = dm.Event(), dm.Event()
        dm.synchronize()
        e0.record(stream)
        for _ in range(args.steps):
            plan.execute(dev, stream=stream.value)
        e1.record(stream)
        e1.synchronize()
        stats = plan.sync()
        plan.close()
        return e0.elapsed_ms(e1) / args.steps, stats, dev["out"].to_host()

    # the project autotuner picks the form-2 settings for this batch first
    tuned, tune_rep = capi.autotune(dev)
    tuned_mixed, tune_rep_mixed = capi.autotune(dev, capi.config(precision=capi.PREC_MIXED))
    ms_per_step, st, fp64_prices = time_steps(tuned)
    # mixed precision on the same batch: FP32 tree passes, FP64 discounting and sums
    mixed_ms, _, mixed_prices = time_steps(tuned_mixed)
    mixed_err = np.abs(mixed_prices - fp64_prices)

    # dominant kernel on its own: the same pipeline launched without the graph,
    # HIP events around the wide-tree kernel on the stream it runs on
    cfg_k = capi.config()
    C.memmove(C.byref(cfg_k), C.byref(tuned), C.sizeof(cfg_k))
    # the wide-tree kernel on its own: serial arrangement whatever the tuner
    # chose (fused, it would include the narrow trees; side by side, it would
    # share the CUs with them)
    cfg_k.flags = (cfg_k.flags & ~(capi.FLAG_FUSED | capi.FLAG_CONCURRENT)) | capi.FLAG_NO_GRAPH
    plan_k = capi.Plan(opts.N, y.N, cfg_k)
    wide, narrow, kern = [], [], []
    for i in range(args.warmup + args.steps):
        plan_k.execute(dev)
        s = plan_k.sync()
        if i >= args.warmup:
            wide.append(s["wide_ms"])
            narrow.append(s["narrow_ms"])
            kern.append(s["kernel_ms"])
    plan_k.close()
    wide_ms = float(np.mean(wide))
    wide_flop = FLOP_PER_NODE_UPDATE * s["wide_node_updates"]
    dominant = "k_price_wide"
    if tuned.form == capi.FORM_LANE:  # the inspector chose form 1: one pricing kernel
        wide_ms, wide_flop, dominant = float(np.mean(kern)), FLOP_PER_NODE_UPDATE * s["node_updates"], "k_price_lane"
    achieved = wide_flop / (wide_ms * 1e-3) / 1e12
    # what the kernel's stencil steps executed, counted by the kernel itself
    # (padding slots included; 4K+1 operations per lane and step in the
    # unit-diagonal form, 5K in the plain one)
    exec_ops = s["wide_step_ops"] if dominant == "k_price_wide" else None
    slots = s["wide_unit_slots"] + s["wide_plain_slots"]

    # the drop-in one-shot call on host arrays (plan creation, H2D, pricing,
    # D2H, plan destruction): the PCIe-inclusive rate, reported beside `value`
    capi.price(opts, y)
    t0 = time.perf_counter()
    for _ in range(3):
        capi.price(opts, y)
    one_shot_s = (time.perf_counter() - t0) / 3

    traffic = None
    tfile = ROOT / "profiles" / "traffic_wide_kernel_skewed_60k.json"
    if tfile.exists():
        traffic = json.loads(tfile.read_text()).get("hbm_bytes_per_launch")

    line = {
        "metric": METRIC,
        "value": opts.N / (ms_per_step * 1e-3),
        "unit": "options/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded skewed 60k batch, 120-knot linear 1%-4% zero curve)",
        "config": {
            "workload": "skewed_60k: 60,000 European zero-coupon-bond options on the Hull-White "
                        "one-factor trinomial tree, FP64; 54,000 small trees (width <= 47, height <= 60) "
                        "and 6,000 large (width 267-1,327, height 360-1,080)",
            "options": opts.N,
            "precision": "fp64",
            "pipeline": "device constants + LPT ordering + wide/narrow persistent pricing kernels, "
                        + ("launched directly" if tuned.flags & capi.FLAG_DIRECT else "one HIP graph"),
            "tuned": tuned.as_dict(),
        },
        "roofline": {
            "bound": "valu",
            "kernel": dominant,
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": traffic,
            "kernel_ms": wide_ms,
            "flop_per_launch": wide_flop,
            "flop_model": "5 FP64 FLOP per stencil node update (mul + 2 fma), node updates counted over "
                          "the active tree region",
            "executed": None if exec_ops is None else {
                "flop_per_launch": exec_ops,
                "achieved": exec_ops / (wide_ms * 1e-3) / 1e12,
                "frac": exec_ops / (wide_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                "unit_form_slots": s["wide_unit_slots"],
                "plain_form_slots": s["wide_plain_slots"],
                "slots_per_counted_update": slots / s["wide_node_updates"],
                "note": "FP64 operations the kernel's stencil steps executed, counted by the kernel (window "
                        "padding included): 4K+1 per lane and step in the unit-diagonal form, 5K in the plain form",
            },
        },
        "detail": {
            "node_updates": st["node_updates"],
            "seq_equivalent_node_updates": st["seq_node_updates"],
            "wide_tasks": st["wide_tasks"],
            "narrow_kernel_ms": float(np.mean(narrow)),
            "pricing_kernels_ms": float(np.mean(kern)),
            "one_shot_host_buffers": {"options_per_s": opts.N / one_shot_s, "ms": one_shot_s * 1e3,
                                      "note": "hwt_price_f64 on host arrays: plan creation, H2D copies, "
                                              "pricing, D2H copy and plan destruction (wall clock)"},
            "total_ms_per_step": ms_per_step,
            "autotune": {"fp64": tune_rep, "mixed": tune_rep_mixed, "mixed_config": tuned_mixed.as_dict()},
            "mixed_precision": {
                "options_per_s": opts.N / (mixed_ms * 1e-3),
                "ms_per_step": mixed_ms,
                "max_abs_err_vs_fp64": float(mixed_err.max()),
                "mean_abs_err_vs_fp64": float(mixed_err.mean()),
                "note": "same batch; FP32 tree passes, FP64 discount factors and compensated FP64 sums; "
                        "error per 100 notional against the FP64 pricer (itself within 1e-10 of seq)",
            },
        },
    }
    if not args.no_reference_sets:
        line["detail"]["reference_sets_65536"] = reference_sets_leg(pkg, capi, args)
        line["detail"]["tiny_trees_1m"] = tiny_trees_leg(pkg, capi, y)
    if not args.no_surrogate:
        line["surrogate"] = surrogate_leg(pkg, y, args)
    if not args.no_cpu_baseline:
        sys.path.insert(0, str(ROOT / "oracle"))
        line["cpu_baseline"] = cpu_baseline(pkg, opts, y, pkg.datasets.uniform_1k())
        cb = line["cpu_baseline"]
        cb["gpu_over_cpu"] = line["value"] / cb["value"]
        cb["gpu_over_cpu_single_thread"] = line["value"] / cb["single_thread_value"]
        mp = line["detail"]["mixed_precision"]
        mp["options_per_s_over_config1_seq"] = mp["options_per_s"] / cb["config1_single_thread_value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
