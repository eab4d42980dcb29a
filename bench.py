#!/usr/bin/env python3
"""Benchmark: full Cu->Cu 500 keV cascades on one MI355X.

Workload: 500 keV Cu ions into a 2 um Cu cube (8.96 g/cm3, Ed 30 eV, Eb 3 eV),
1e6 histories, FP64, every recoil followed (mytrim_amd.configs.c2_cu_cu).  A
step is one batch of histories (one persistent-kernel launch); the K timed
steps together cover all --histories histories (distinct random streams per
history), the W warm-up steps run on extra histories.  By default the whole
job is one launch: every launch ends with a tail in which few lanes still
hold a flight, so fewer, larger launches keep more of the chip busy.  The launch shape is
autotuned on a sample first (outside the timed region).

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
FP32_VALU_PEAK_TFLOPS = 157.3
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


def train_surrogate(eng, gpu, prob, n_train: int, epochs: int):
    opts = gpu.Engine.surrogate_opts(prob, n_train=n_train, epochs=epochs, bf16=True)
    t0 = time.perf_counter()
    params, loss = eng.surrogate_train(opts, gpu.Engine.surrogate_init(1))
    t_train = time.perf_counter() - t0
    eng.surrogate_install(opts, params)
    return {"train_s": t_train, "final_loss": float(loss[-1]), "n_train": n_train, "epochs": epochs,
            "domain": {k: v for k, v in eng.surrogate_norm(opts).items() if k in ("eps_lo", "eps_hi", "b_hi")}}


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
            "sample_rate": rate}


def timed(eng, prob, arith, n_hist, steps):
    bounds = [n_hist * i // steps for i in range(steps + 1)]
    coll, kms, res = 0, 0.0, []
    t0 = time.perf_counter()
    for i in range(steps):
        r = eng.run(prob.with_(hist0=bounds[i], n_hist=bounds[i + 1] - bounds[i]), arith=arith)
        coll += r["n_coll"]
        kms += r["kernel_ms"]
        res.append(r)
    return coll, time.perf_counter() - t0, kms, res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--histories", type=int, default=1000000)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--autotune-sample", type=int, default=10000,
                    help="least histories in the tuning sample (0 = default or given launch shape)")
    ap.add_argument("--block", type=int, default=0, help="fixed launch shape (with --autotune-sample 0)")
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--stack-cap", type=int, default=0)
    ap.add_argument("--variant-histories", type=int, default=100000,
                    help="histories for the FP32 and FP32+surrogate lines (0 = skip)")
    ap.add_argument("--surrogate-train", type=int, default=4_000_000)
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
    if args.block or args.blocks_per_cu or args.stack_cap:
        eng.tune(block=args.block, blocks_per_cu=args.blocks_per_cu, stack_cap=args.stack_cap)
        tuning = {"block": args.block, "blocks_per_cu": args.blocks_per_cu, "stack_cap": args.stack_cap,
                  "fixed": True}
    elif args.autotune_sample > 0:
        # the tuning sample holds at least --autotune-sample histories and about
        # 3e9 collisions, at most one timed step: small cascades need more
        # histories to reach the queue occupancy of a full launch, and mixed
        # workloads a sample close to a launch's size
        pilot = eng.run(prob.with_(hist0=3 * args.histories, n_hist=2000))
        per_hist = max(pilot["n_coll"] / 2000.0, 1.0)
        n_sample = int(min(max(args.autotune_sample, 3e9 / per_hist), max(args.histories // max(args.steps, 1), 1)))
        best, rate = eng.autotune(prob.with_(hist0=2 * args.histories, n_hist=n_sample), sample=n_sample)
        tuning = {"block": best.block, "blocks_per_cu": best.blocks_per_cu, "stack_cap": best.stack_cap,
                  "sample_histories": n_sample, "sample_rate": rate}
    wsize = max(1, min(20000, args.histories // max(args.steps, 1)))
    for i in range(args.warmup):
        eng.run(prob.with_(hist0=args.histories + i * wsize, n_hist=wsize))
    coll, wall, kernel_ms, steps = timed(eng, prob, gpu.MT_FP64, args.histories, args.steps)

    variants = None
    if args.variant_histories > 0:
        nv = min(args.variant_histories, args.histories)
        c64, w64, k64, r64 = timed(eng, prob, gpu.MT_FP64, nv, 1)
        # each precision runs at its own autotuned launch shape (the FP32
        # kernels need fewer registers, so more waves fit), tuned on a sample
        # the size of the variant's own launch
        sample = prob.with_(hist0=2 * args.histories, n_hist=nv)
        t32 = _tuned(eng, sample, gpu.MT_FP32, args.autotune_sample)
        c32, w32, k32, r32 = timed(eng, prob, gpu.MT_FP32, nv, 1)
        sur = train_surrogate(eng, gpu, prob, args.surrogate_train, args.surrogate_epochs)
        tsur = _tuned(eng, sample, gpu.MT_FP32_SURROGATE, args.autotune_sample)
        cs, ws, ks, rs = timed(eng, prob, gpu.MT_FP32_SURROGATE, nv, 1)
        variants = {
            "histories": nv,
            "fp64": {"collisions_per_s": c64 / w64, "kernel_ms": k64},
            "fp32": {"collisions_per_s": c32 / w32, "kernel_ms": k32, "speedup_vs_fp64": (c32 / w32) / (c64 / w64),
                     "tuning": t32, "tally_rel_err_vs_fp64": tally_errors(r64[0], r32[0])},
            "fp32_surrogate": {"collisions_per_s": cs / ws, "kernel_ms": ks, "tuning": tsur,
                               "speedup_vs_fp64": (cs / ws) / (c64 / w64),
                               "surrogate_share": rs[0]["surrogate"] / max(rs[0]["n_coll"], 1),
                               "tally_rel_err_vs_fp64": tally_errors(r64[0], rs[0]),
                               "training": sur},
        }
    eng.close()

    w = work_per_collision(args.config)
    n_launch = len(steps)
    avg_ms = kernel_ms / n_launch
    flop_per_launch = coll * w["flop_eq"] / n_launch
    achieved = flop_per_launch / (avg_ms * 1e-3) / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_%s.json" % args.config)
    if os.path.exists(tpath):
        with open(tpath) as f:
            bpc = json.load(f).get("bytes_per_collision")
        if bpc:
            traffic = bpc * coll / n_launch
    s0 = steps[0]
    out = {
        "metric": METRIC,
        "value": coll / wall,
        "unit": "collisions/s",
        "histories_per_s": args.histories / wall,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": prob.name, "histories": args.histories,
                   "cascades": "full (every recoil followed)", "arith": "FP64",
                   "collisions": coll, "collisions_per_history": coll / args.histories,
                   "grid": s0["grid"], "block": s0["block"], "waves": s0["waves"],
                   "queue_peak": max(r["queue_peak"] for r in steps), "tuning": tuning},
        "roofline": {"bound": "mfma", "pipe": "FP64 VALU (the FP64 vector and matrix peaks are equal)",
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": "cascade_kernel<0,0,0,false>", "launches": n_launch,
                     "avg_kernel_ms": avg_ms, "flop_eq_per_collision": w["flop_eq"],
                     "flop_eq_per_launch": flop_per_launch},
        "variants": variants,
    }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        n_cpu = args.cpu_histories
        if n_cpu <= 0:
            # about 15 s of CPU work at ~2.6e6 collisions/s per core
            n_cpu = max(16, int(15.0 * 2.6e6 * threads / max(coll / args.histories, 1.0)))
        out["cpu_baseline"] = cpu_baseline(prob, n_cpu, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
