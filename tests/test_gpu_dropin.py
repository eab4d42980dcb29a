"""The reference's own regression test (tests/uo2 of MyTRIM: fission fragments
in UO2 with Xe bubbles, MYTRIM_SEED=39172, r = 10 A, Cbf = 0.1, one fission
event) run through the C++ drop-in API: every TrimBase::trim() flight
executes on the GPU with the SimconfType stream and the hooks are replayed
on the host.  Compared like the reference's tests/csv_diff.sh: column by
column, absolute tolerance 1e-4.

Two callers run it:
* tests/native/fission_bubbles.cpp, this build's own driver of the scenario;
* the reference's unchanged apps/mytrim_uo2.C, compiled against the drop-in
  headers and libmytrim.so where the reference sources were mounted at build
  time (oracle/_ref/mytrim_uo2_dropin, oracle/Makefile); skipped without it.
"""
import os
import subprocess
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OWN = os.path.join(ROOT, "mytrim_amd", "host", "_build", "fission_bubbles")
REF_CALLER = os.path.join(ROOT, "oracle", "_ref", "mytrim_uo2_dropin")
GOLD = os.path.join(ROOT, "tests", "golden", "uo2")


def run_and_compare(exe, tmp_path):
    env = dict(os.environ, MYTRIM_SEED="39172")
    t0 = time.perf_counter()
    r = subprocess.run([exe, "out", "10", "0.1", "1"], cwd=tmp_path, env=env, capture_output=True,
                       text=True, timeout=600)
    dt = time.perf_counter() - t0
    assert r.returncode == 0, r.stderr[-2000:]
    for name in ("out.Erec", "out.clcoor", "out.dist"):
        mine = np.loadtxt(tmp_path / name, ndmin=2)
        gold = np.loadtxt(os.path.join(GOLD, name), ndmin=2)
        assert mine.shape == gold.shape, name
        assert np.max(np.abs(mine - gold)) <= 1e-4, name
    return r.stdout, dt


def test_uo2_gold_through_gpu_flights(tmp_path):
    assert os.path.exists(OWN), "build the C++ API first"
    out, dt = run_and_compare(OWN, tmp_path)
    print("fission_bubbles: %s  %.2f s" % (out.strip(), dt))


@pytest.mark.skipif(not os.path.exists(REF_CALLER), reason="unchanged reference caller not built here")
def test_unchanged_reference_caller_reproduces_gold(tmp_path):
    out, dt = run_and_compare(REF_CALLER, tmp_path)
    print("apps/mytrim_uo2.C on the drop-in: %.2f s" % dt)
