"""The autotuner's launch-rate projection (tests/native/tune_check.cpp over
mt_tune.h, the header mt_autotune uses): a launch is a ramp, a steady phase
and a drain; the rate projected for a larger batch keeps the sample's ramp
and drain and runs the rest at the steady rate."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_projected_rate(tmp_path):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "tune_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "mytrim_amd", "csrc"),
                    "-I" + os.path.join(ROOT, "include"), "-o", exe,
                    os.path.join(HERE, "native", "tune_check.cpp")], check=True, capture_output=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    lines = {line.split()[0]: line.split()[1:] for line in out.strip().splitlines()}
    r = {k: [float(x) for x in v] for k, v in lines.items() if not k.startswith(("pick", "final"))}
    whole = r["same"][1]
    # the sample's own size, or a smaller batch: the sample's kernel-time rate
    assert r["same"][0] == pytest.approx(whole, rel=1e-12)
    assert r["smaller"][0] == pytest.approx(whole, rel=1e-12)
    # a batch of 1e6 histories: ramp + the rest at the steady rate + drain
    assert r["batch"][0] == pytest.approx(r["batch"][1], rel=1e-6)
    assert whole < r["batch"][0] < 2.0e10
    # a batch so large that ramp and drain vanish: the steady rate
    assert r["huge"][0] == pytest.approx(2.0e10, rel=1e-6)
    # missing or disordered stamps and a short steady phase fall back to kernel time
    for k in ("nostamp", "order", "short"):
        assert r[k][0] == pytest.approx(whole, rel=1e-12), k
    # picks: the fastest; the deepest stack within 2 % of it; none without a rate
    assert lines["pick_fastest"] == ["0"]  # 256 x 2 with 128 entries: within 2 % of 96, deeper
    assert lines["pick_deeper"] == ["1"]   # 192 entries, 1.5 % behind the fastest
    assert lines["pick_none"] == ["-1"]
    # finalists by occupancy, ascending: 8 waves per CU its pick (the deepest
    # stack within 2 %) and its runner-up of another shape; 12 its pick (the
    # other shape is over 10 % behind the fastest); 4 none
    assert lines["final_runner"] == ["3", "256:2:192:0", "256:2:128:0", "256:3:96:0"]
    assert lines["final_pick"] == ["2", "256:2:192:0", "256:3:96:0"]
