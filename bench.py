#!/usr/bin/env python3
"""Benchmark: full Cu->Cu 500 keV cascades on one MI355X.

Workload: 500 keV Cu ions into a 2 um Cu cube (8.96 g/cm3, Ed 30 eV, Eb 3 eV),
1e6 histories, FP64, every recoil followed (mytrim_amd.configs.c2_cu_cu).  A
step is one pass of the persistent cascade kernel over one batch of
--histories histories (default 1e6, the workload's size): step i runs
histories [i H, (i + 1) H), so every step has its own random streams.  The W
warm-up steps run smaller batches on other histories, untimed.  The launch
shape is autotuned on a sample first (outside the timed region).

Prints ONE JSON line: collisions/s (value) and histories/s of the whole job,
the same workload in FP32 and FP32 + learned scattering surrogate on a
sub-batch, the roofline of the cascade kernel (FP64 work per collision,
frozen by counting the operations of the CPU restatement, over the kernel time
measured with HIP events on its stream), and the CPU baseline: the
reference's algorithm restated in C++ (std::mt19937 draws), timed on this
host's cores on a bounded sample of the same workload.

--copy-calibration runs only a device-to-device copy of known size (the
reference kernel for scaling FETCH_SIZE / WRITE_SIZE counters to bytes).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix alike)
# MI355X FP32 vector peak without packed FP32 instructions: 157.3 TFLOP/s
# counts the packed (two-wide) FMA; the kernels are built with packed FP32
# disabled, so one FP32 FMA per lane per clock, the same rate as FP64
FP32_UNPACKED_PEAK_TFLOPS = 78.6
CPU_CORES_MAX = 16        # the host-core share of one GPU on the benchmark machines
METRIC = "collisions/sec + ion histories/sec on 1 MI355X (plain FP64 vs FP32+surrogate)"


def work_per_collision(config: str) -> dict:
    path = os.path.join(ROOT, "mytrim_amd", "data", "work_per_collision.json")
    with open(path) as f:
        return json.load(f)[config]


def cpu_baseline(prob, n_hist: int, threads: int) -> dict:
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    sample = prob.with_(n_hist=n_hist, hist0=0)
    r = O.run(sample, rng=0, nthreads=threads)
    return {"value": r["n_coll"] / r["seconds"], "unit": "collisions/s", "cores": threads,
            "kind": "port",
            "histories_per_s": n_hist / r["seconds"],
            "sample": "%d histories of the same workload (%d collisions, %.1f s), FP64, "
                      "std::mt19937 streams, one thread per core" % (n_hist, r["n_coll"], r["seconds"]),
            "cpu": _cpu_model()}


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def copy_calibration(mib: int, reps: int):
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "reps": reps}), flush=True)


def train_surrogate(eng, gpu, prob, n_train: int, epochs: int, **kw):
    """Train on the device and install; kw: inputs, sampling, eps_min, eps_max."""
    opts = gpu.Engine.surrogate_opts(prob, n_train=n_train, epochs=epochs, bf16=True, **kw)
    t0 = time.perf_counter()
    params, loss = eng.surrogate_train(opts, gpu.Engine.surrogate_init(1))
    t_train = time.perf_counter() - t0
    eng.surrogate_install(opts, params)
    norm = eng.surrogate_norm(opts)
    return {"train_s": t_train, "final_loss": float(loss[-1]), "n_train": n_train, "epochs": epochs,
            "inputs": opts.inputs, "sampling": ["as the cascades draw", "rectangle"][opts.sampling],
            "domain": {k: v for k, v in norm.items() if k in ("eps_lo", "eps_hi", "b_hi")},
            "accuracy": surrogate_accuracy(eng, gpu, norm),
            "inference_ablation": inference_ablation(eng, gpu, norm)}


def inference_ablation(eng, gpu, norm, n=1 << 21):
    """Time per scattering evaluation of the FP32 MAGIC solve and of the installed
    network with layer 2 in FP32 FMAs, BF16 dot products and BF16 matrix-core
    products (mt_surrogate_ablate), on the same pairs of the surrogate's domain
    (eps log-uniform up to min(eps_hi, 10), b uniform), each with its |d ln s2|
    against the device's FP64 solve."""
    import numpy as np
    rng = np.random.default_rng(11)
    hi = min(norm["eps_hi"], 10.0)
    eps = np.exp(rng.uniform(np.log(norm["eps_lo"]), np.log(hi), n))
    b = rng.uniform(0.0, norm["b_hi"], n)
    ref, _, _, _ = eng.scatter(eps, b, 0)
    out = {"pairs": n}
    t_magic = None
    for mode, name in gpu.Engine.ABLATE_MODES.items():
        s2, ms = eng.surrogate_ablate(mode, eps, b, reps=5)
        ok = (s2 > 0) & (ref > 0)
        d = np.abs(np.log(s2[ok]) - np.log(ref[ok]))
        t_magic = ms if mode == 0 else t_magic
        out[name] = {"evaluations_per_s": n / (ms * 1e-3), "ns_per_evaluation": ms * 1e6 / n,
                     "time_vs_magic_fp32": ms / t_magic, "median_abs_dln_s2": float(np.median(d)),
                     "p99_abs_dln_s2": float(np.quantile(d, 0.99))}
    return out


def surrogate_accuracy(eng, gpu, norm, n=200000):
    """|d ln s2| of the installed surrogate's scattering step against the
    device's FP64 solve, on eps log-uniform and b uniform over the part of the
    domain the surrogate answers (eps <= 10)."""
    import numpy as np
    rng = np.random.default_rng(5)
    hi = min(norm["eps_hi"], 10.0)
    eps = np.exp(rng.uniform(np.log(norm["eps_lo"]), np.log(hi), n))
    b = rng.uniform(0.0, norm["b_hi"], n)
    s2s, _, _, its = eng.scatter(eps, b, 0, arith=gpu.MT_FP32_SURROGATE)
    s2d, _, _, _ = eng.scatter(eps, b, 0)
    ok = (its < 0) & (s2d > 0) & (s2s > 0)
    d = np.abs(np.log(s2s[ok]) - np.log(s2d[ok]))
    return {"pairs": int(ok.sum()), "median_abs_dln_s2": float(np.median(d)),
            "p99_abs_dln_s2": float(np.quantile(d, 0.99))}


def tally_errors(ref, r):
    """Relative difference of each tally from the FP64 run on the same histories."""
    def rel(a, b):
        return (b - a) / a if a else 0.0
    out = {k: rel(ref[k], r[k]) for k in ("n_coll", "n_flight", "vac", "repl", "disp", "Eel", "Enuc")}
    out["range_mean"] = rel(ref["range_sum"] / max(ref["range_n"], 1), r["range_sum"] / max(r["range_n"], 1))
    for h in ("range_hist", "vac_hist"):  # total-variation distance of the histograms
        a, b = ref[h], r[h]
        out[h + "_tv"] = 0.5 * float(abs(a / max(a.sum(), 1) - b / max(b.sum(), 1)).sum())
    return out


def _tuned(eng, sample, arith, n_sample):
    """Autotune the launch shape for one precision (kept for the next runs)."""
    if n_sample <= 0:
        return None
    best, rate = eng.autotune(sample, sample=sample.n_hist, arith=arith)
    return {"block": best.block, "blocks_per_cu": best.blocks_per_cu, "stack_cap": best.stack_cap,
            "primary_waves64": getattr(best, "primary_waves64", 0), "lds_hist": best.lds_hist, "sample_rate": rate}


def step_problem(prob, i, n_hist):
    """Step i: histories [i n, (i + 1) n) (their own random streams); per-history
    primaries (the mixed workload) repeat the first n rows."""
    return prob.batch(0, n_hist).with_(hist0=i * n_hist)


def timed(eng, prob, arith, n_hist, steps, first=0):
    coll, kms, res = 0, 0.0, []
    t0 = time.perf_counter()
    for i in range(steps):
        r = eng.run(step_problem(prob, first + i, n_hist), arith=arith)
        coll += r["n_coll"]
        kms += r["kernel_ms"]
        res.append(r)
    return coll, time.perf_counter() - t0, kms, res


def roofline(coll, launches, kernel_ms, flop_eq, peak, pipe, kernel, traffic_bpc=None):
    """Achieved TFLOP/s of one kernel: algorithmic work per launch (flop-eq
    per collision x collisions) over the average launch duration."""
    avg_ms = kernel_ms / launches
    flop_per_launch = coll * flop_eq / launches
    achieved = flop_per_launch / (avg_ms * 1e-3) / 1e12
    return {"bound": "valu", "pipe": pipe, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": traffic_bpc * coll / launches if traffic_bpc else None,
            "kernel": kernel, "launches": launches, "avg_kernel_ms": avg_ms,
            "flop_eq_per_collision": flop_eq, "flop_eq_per_launch": flop_per_launch}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--histories", type=int, default=1000000, help="histories per step")
    ap.add_argument("--config", default="c2")
    ap.add_argument("--autotune-sample", type=int, default=10000,
                    help="least histories in the tuning sample (0 = default or given launch shape)")
    ap.add_argument("--block", type=int, default=0, help="fixed launch shape (with --autotune-sample 0)")
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--stack-cap", type=int, default=0)
    ap.add_argument("--primary-waves64", type=int, default=0,
                    help="feeder waves per 64 with a fixed launch shape (0: no wave roles)")
    ap.add_argument("--lds-hist", type=int, default=0,
                    help="depth bins in LDS with a fixed launch shape (0: as many as fit, -1: none)")
    ap.add_argument("--variant-histories", type=int, default=100000,
                    help="histories for the FP32 and FP32+surrogate lines (0 = skip)")
    ap.add_argument("--surrogate-train", type=int, default=4_000_000)
    ap.add_argument("--surrogate-train-configured", type=int, default=10_000_000)
    ap.add_argument("--surrogate-epochs", type=int, default=60)
    ap.add_argument("--cpu-histories", type=int, default=0, help="0 = size to ~15 s")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X. Histories are "
              "independent, so N GPUs would each take a contiguous range of history ids; "
              "no multi-GPU run is made here.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration, 5)
        return

    from mytrim_amd import configs, gpu
    build = {"c1": configs.c1_he_fe, "c2": configs.c2_cu_cu, "c3": configs.c3_ag_layers,
             "c4": configs.c4_mixed_uo2, "c5": configs.c5_c_w}[args.config]
    prob = build(args.histories)
    eng = gpu.Engine(0)
    eng.set_target(prob)
    tuning = None
    if args.block or args.blocks_per_cu or args.stack_cap or args.primary_waves64 or args.lds_hist:
        eng.tune(block=args.block, blocks_per_cu=args.blocks_per_cu, stack_cap=args.stack_cap,
                 primary_waves64=args.primary_waves64, lds_hist=args.lds_hist)
        tuning = {"block": args.block, "blocks_per_cu": args.blocks_per_cu, "stack_cap": args.stack_cap,
                  "primary_waves64": args.primary_waves64, "lds_hist": args.lds_hist, "fixed": True}
    elif args.autotune_sample > 0:
        # the tuning sample holds at least --autotune-sample histories and about
        # 3e9 collisions, at most one timed step: small cascades need more
        # histories to reach the queue occupancy of a full launch, and mixed
        # workloads a sample close to a launch's size
        # tuning histories come after every timed and warm-up step's
        spare = (args.steps + args.warmup + 2) * args.histories
        pilot = eng.run(step_problem(prob, 0, 2000).with_(hist0=spare))
        per_hist = max(pilot["n_coll"] / 2000.0, 1.0)
        n_sample = int(min(max(args.autotune_sample, 3e9 / per_hist), args.histories))
        # tuned for a launch of one step's histories: the tuner times its
        # finalists again on up to four times the sample
        n_final = int(min(4 * n_sample, args.histories))
        best, rate = eng.autotune(step_problem(prob, 0, args.histories).with_(hist0=spare + 2000), sample=n_sample)
        tuning = {"block": best.block, "blocks_per_cu": best.blocks_per_cu, "stack_cap": best.stack_cap,
                  "primary_waves64": getattr(best, "primary_waves64", 0), "lds_hist": best.lds_hist,
                  "sample_histories": n_sample, "final_histories": n_final, "sample_rate": rate}
    wsize = max(1, min(20000, args.histories))
    for i in range(args.warmup):
        eng.run(step_problem(prob, 0, wsize).with_(hist0=args.steps * args.histories + i * wsize))
    coll, wall, kernel_ms, steps = timed(eng, prob, gpu.MT_FP64, args.histories, args.steps)

    variants = None
    if args.variant_histories > 0:
        nv = min(args.variant_histories, args.histories)
        w_eq = work_per_collision(args.config)["flop_eq"]
        c64, w64, k64, r64 = timed(eng, prob, gpu.MT_FP64, nv, 1)
        # each precision runs at its own autotuned launch shape (the FP32
        # kernels need fewer registers, so more waves fit), tuned on a sample
        # the size of the variant's own launch
        spare = (args.steps + args.warmup + 4) * args.histories
        sample = step_problem(prob, 0, nv).with_(hist0=spare)
        t32 = _tuned(eng, sample, gpu.MT_FP32, args.autotune_sample)
        c32, w32, k32, r32 = timed(eng, prob, gpu.MT_FP32, nv, 1)
        # the configured network: 2 -> 64 -> 64 -> 1 on 1e7 (log eps, b) pairs of a
        # rectangle, eps in [1e-4, 1e4]
        surc = train_surrogate(eng, gpu, prob, args.surrogate_train_configured, args.surrogate_epochs // 2,
                               inputs=2, sampling=1, eps_min=1e-4, eps_max=1e4)
        csc, wsc, ksc, rsc = timed(eng, prob, gpu.MT_FP32_SURROGATE, nv, 1)
        # this build's network: 3 inputs, trained on the pairs the cascades draw
        sur = train_surrogate(eng, gpu, prob, args.surrogate_train, args.surrogate_epochs)
        tsur = _tuned(eng, sample, gpu.MT_FP32_SURROGATE, args.autotune_sample)
        cs, ws, ks, rs = timed(eng, prob, gpu.MT_FP32_SURROGATE, nv, 1)
        pipe32 = "FP32 VALU, packed FP32 disabled (one FMA per lane per clock: the FP64 rate)"
        variants = {
            "histories": nv,
            "fp64": {"collisions_per_s": c64 / w64, "kernel_ms": k64,
                     "roofline": roofline(c64, 1, k64, w_eq, FP64_PEAK_TFLOPS, "FP64 VALU",
                                          "cascade_kernel<0,0,0,false>")},
            "fp32": {"collisions_per_s": c32 / w32, "kernel_ms": k32, "speedup_vs_fp64": (c32 / w32) / (c64 / w64),
                     "tuning": t32, "tally_rel_err_vs_fp64": tally_errors(r64[0], r32[0]),
                     "roofline": roofline(c32, 1, k32, w_eq, FP32_UNPACKED_PEAK_TFLOPS, pipe32,
                                          "cascade_kernel<1,0,0,false>")},
            "fp32_surrogate": {"collisions_per_s": cs / ws, "kernel_ms": ks, "tuning": tsur,
                               "speedup_vs_fp64": (cs / ws) / (c64 / w64),
                               "surrogate_share": rs[0]["surrogate"] / max(rs[0]["n_coll"], 1),
                               "tally_rel_err_vs_fp64": tally_errors(r64[0], rs[0]),
                               "roofline": roofline(cs, 1, ks, w_eq, FP32_UNPACKED_PEAK_TFLOPS, pipe32,
                                                    "cascade_kernel<2,0,0,false>"),
                               "training": sur},
            "fp32_surrogate_configured": {
                "collisions_per_s": csc / wsc, "kernel_ms": ksc, "speedup_vs_fp64": (csc / wsc) / (c64 / w64),
                "surrogate_share": rsc[0]["surrogate"] / max(rsc[0]["n_coll"], 1),
                "tally_rel_err_vs_fp64": tally_errors(r64[0], rsc[0]),
                "roofline": roofline(csc, 1, ksc, w_eq, FP32_UNPACKED_PEAK_TFLOPS, pipe32,
                                     "cascade_kernel<2,0,0,false>"),
                "training": surc},
        }
    eng.close()

    w = work_per_collision(args.config)
    bpc = None
    tpath = os.path.join(ROOT, "profiles", "traffic_%s.json" % args.config)
    if os.path.exists(tpath):
        with open(tpath) as f:
            bpc = json.load(f).get("bytes_per_collision")
    s0 = steps[0]
    out = {
        "metric": METRIC,
        "value": coll / wall,
        "unit": "collisions/s",
        "histories_per_s": args.histories * args.steps / wall,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": prob.name, "histories_per_step": args.histories,
                   "histories": args.histories * args.steps,
                   "cascades": "full (every recoil followed)", "arith": "FP64",
                   "collisions": coll, "collisions_per_history": coll / (args.histories * args.steps),
                   "grid": s0["grid"], "block": s0["block"], "waves": s0["waves"],
                   "queue_peak": max(r["queue_peak"] for r in steps), "tuning": tuning},
        "roofline": roofline(coll, len(steps), kernel_ms, w["flop_eq"], FP64_PEAK_TFLOPS,
                             "FP64 VALU (no matrix work: the kernel issues no MFMA)",
                             "cascade_kernel<0,0,0,false>", bpc),
        "variants": variants,
    }
    # the waves' mean shader clock over the timed launches: the published FP64
    # peak assumes the 2400 MHz maximum
    sclk = sum(r["sclk_mhz"] for r in steps) / len(steps)
    out["roofline"]["sclk_mhz"] = sclk
    out["roofline"]["peak_at_sclk"] = FP64_PEAK_TFLOPS * sclk / 2400.0
    if not args.no_cpu:
        # one thread per core of the GPU's host-core share (at most 16 worker
        # threads per GPU on the benchmark machines, whatever os.cpu_count()
        # reports there)
        threads = min(CPU_CORES_MAX, os.cpu_count() or 1)
        n_cpu = args.cpu_histories
        if n_cpu <= 0:
            # about 15 s of CPU work at ~2.6e6 collisions/s per core
            n_cpu = max(16, int(15.0 * 2.6e6 * threads / max(coll / (args.histories * args.steps), 1.0)))
        out["cpu_baseline"] = cpu_baseline(prob, n_cpu, threads)
        out["cpu_baseline"]["os_cpu_count"] = os.cpu_count()
        out["cpu_baseline"]["cores_note"] = ("one std::thread per core, capped at the %d host cores one GPU may use "
                                             "on the benchmark machines" % CPU_CORES_MAX)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
