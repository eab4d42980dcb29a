#!/usr/bin/env python3
"""Benchmark: full Cu->Cu 500 keV cascades on one MI355X.

Workload: 500 keV Cu ions into a 2 um Cu cube (8.96 g/cm3, Ed 30 eV, Eb 3 eV),
1e6 histories, FP64, every recoil followed (mytrim_amd.configs.c2_cu_cu).  A
step is one batch of histories; the K timed steps together cover all
--histories histories (distinct random streams per history), the W warm-up
steps run on extra histories.

Prints ONE JSON line: collisions/s (value) and histories/s of the whole job,
the roofline of the cascade kernel (FP64 VALU work per collision, frozen by
counting the operations of the CPU restatement, over the kernel time measured
with HIP events on its stream), and the CPU baseline: the reference's
algorithm restated in C++ (std::mt19937 draws), timed on this host's cores on
a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_VALU_PEAK_TFLOPS = 78.6      # MI355X dense FP64 vector
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--histories", type=int, default=1000000)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--cpu-histories", type=int, default=0, help="0 = size to ~15 s")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X; the "
              "histories would be split across GPUs as described in DESIGN.md", file=sys.stderr)
        sys.exit(2)

    from mytrim_amd import configs, gpu
    build = {"c1": configs.c1_he_fe, "c2": configs.c2_cu_cu, "c3": configs.c3_ag_layers,
             "c5": configs.c5_c_w}[args.config]
    prob = build(args.histories)
    eng = gpu.Engine(0)
    eng.set_target(prob)
    # warm-up on histories beyond the timed range
    wsize = max(1, min(20000, args.histories // max(args.steps, 1)))
    for i in range(args.warmup):
        eng.run(prob.with_(hist0=args.histories + i * wsize, n_hist=wsize))
    # timed steps: K batches covering [0, histories)
    bounds = [args.histories * i // args.steps for i in range(args.steps + 1)]
    coll = 0
    kernel_ms = 0.0
    steps = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        r = eng.run(prob.with_(hist0=bounds[i], n_hist=bounds[i + 1] - bounds[i]))
        coll += r["n_coll"]
        kernel_ms += r["kernel_ms"]
        steps.append(r)
    wall = time.perf_counter() - t0
    eng.close()

    w = work_per_collision(args.config)
    achieved = coll * w["flop_eq"] / (kernel_ms * 1e-3) / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_%s.json" % args.config)
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_collision")
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
        "config": {"workload": prob.name, "histories": args.histories, "ion": "Cu 63.546 u, 500 keV",
                   "target": "Cu 8.96 g/cm3, 2 um cube, Ed 30 eV, Eb 3 eV",
                   "cascades": "full (every recoil followed)", "arith": "FP64",
                   "collisions": coll, "collisions_per_history": coll / args.histories,
                   "grid": steps[0]["grid"], "block": steps[0]["block"], "waves": steps[0]["waves"]},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_VALU_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_VALU_PEAK_TFLOPS,
                     "traffic": traffic,
                     "kernel": "cascade_fp64", "kernel_ms": kernel_ms,
                     "flop_eq_per_collision": w["flop_eq"]},
    }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        n_cpu = args.cpu_histories
        if n_cpu <= 0:
            # about 15 s of CPU work at ~1.2e6 collisions/s per core
            n_cpu = max(16, int(15.0 * 1.2e6 * threads / max(coll / args.histories, 1.0)))
        out["cpu_baseline"] = cpu_baseline(prob, n_cpu, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
