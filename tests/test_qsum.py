"""The queue checksum detects torn entries (tests/native/qsum_check.cpp over
mt_qsum.h, the header the cascade kernel includes): a reader that loaded some
fields of the slot's previous lap must see a mismatch and load again."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_torn_entries_are_detected(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "qsum_check")
    subprocess.run([hipcc, "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "mytrim_amd", "csrc"), "-o", exe,
                    os.path.join(HERE, "native", "qsum_check.cpp")], check=True, capture_output=True)
    n = 1_500_000
    out = subprocess.run([exe, str(n), "11"], check=True, capture_output=True, text=True).stdout
    r = {k: int(v) for k, v in (line.split() for line in out.strip().splitlines())}
    assert r["same_ok"] == n  # an entry read whole sums to the writer's value
    assert r["trials"] == 2 * n
    # a 32-bit check misses about one mixture in 4.3e9; 3e6 trials expect 0.0007
    assert r["missed"] == 0, r
