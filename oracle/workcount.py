#!/usr/bin/env python3
"""ORACLE — test infrastructure.  Freezes the algorithmic FP64 work per
collision of the reference's collision loop for each benchmark workload
(counting scalar over the oracle's stopping and scattering, statement count
of the rest of the loop; see oracle/src/runs.cpp, WorkHooks) and writes
mytrim_amd/data/work_per_collision.json, which bench.py reads for the
roofline.  Usage: python oracle/workcount.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle_lib as O  # noqa: E402
from mytrim_amd import configs  # noqa: E402


def main():
    samples = {"c1": configs.c1_he_fe(2000), "c2": configs.c2_cu_cu(40),
               "c3": configs.c3_ag_layers(40), "c4": configs.c4_mixed_uo2(400),
               "c5": configs.c5_c_w(200)}
    out = {}
    for name, prob in samples.items():
        w = O.workcount(prob)
        w["sample_histories"] = prob.n_hist
        w["workload"] = prob.name
        out[name] = w
        print(name, "%.1f flop-eq/collision over %d collisions" % (w["flop_eq"], w["collisions"]))
    out["_weights"] = ("add/sub/mul/cmp/min/max 1, div 10, sqrt 10, exp 20, log 20, pow 40, "
                       "cbrt 30, sin 25, cos 25, atan2 40; (ion, target) constants not counted; "
                       "uniform draws counted apart")
    path = os.path.join(ROOT, "mytrim_amd", "data", "work_per_collision.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
