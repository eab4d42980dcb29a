"""The reference's own regression test (tests/uo2 of MyTRIM: mytrim_uo2 with
MYTRIM_SEED=39172, r = 10 A, Cbf = 0.1, one fission event) run through the
C++ drop-in API: every TrimBase::trim() flight executes on the GPU with the
SimconfType stream and the hooks replayed on the host.  Compared like the
reference's tests/csv_diff.sh: column by column, absolute tolerance 1e-4."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mytrim_amd", "host", "_build", "mytrim_uo2")
GOLD = os.path.join(ROOT, "tests", "golden", "uo2")


def test_uo2_gold_through_gpu_flights(tmp_path):
    assert os.path.exists(EXE), "build the C++ API first"
    env = dict(os.environ, MYTRIM_SEED="39172")
    r = subprocess.run([EXE, "out", "10", "0.1", "1"], cwd=tmp_path, env=env, capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    for name in ("out.Erec", "out.clcoor", "out.dist"):
        mine = np.loadtxt(tmp_path / name, ndmin=2)
        gold = np.loadtxt(os.path.join(GOLD, name), ndmin=2)
        assert mine.shape == gold.shape, name
        assert np.max(np.abs(mine - gold)) <= 1e-4, name
