"""The reference's own callers build unchanged against the drop-in C++ API.

Every research driver of the reference (apps/mytrim_*.C) and its output
modules (apps/src/*.C) is compiled, as it is, against this build's headers
(mytrim_amd/host/include, the reference's class and member names) and linked
against libmytrim.so, whose TrimBase::trim runs each flight on the GPU.
Only compiling and linking happens here (no GPU); tests/test_gpu_dropin.py
runs the unchanged apps/mytrim_uo2.C on the GPU against the reference's gold
files.  Needs the reference sources (skipped where they are not mounted);
apps/runmytrim.C and apps/runstopping.C need jsoncpp, which is not
installed, so their drop-in twins are mytrim_amd/host/apps/*.cpp."""
import glob
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
HOST = os.path.join(ROOT, "mytrim_amd", "host")
LIBDIR = os.path.join(HOST, "_build")
GPULIB = os.path.join(ROOT, "mytrim_amd", "_build")

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "apps")),
                                reason="reference sources not mounted")

DRIVERS = sorted(glob.glob(os.path.join(REF, "apps", "mytrim_*.C")))
MODULES = sorted(glob.glob(os.path.join(REF, "apps", "src", "*.C")))


def _compile(src, out, link):
    inc = ["-I" + os.path.join(HOST, "include"), "-I" + os.path.join(HOST, "apps", "include")]
    cmd = ["g++", "-std=c++11", "-O0", "-w"] + inc
    if link:
        cmd += ["-o", out, src, "-L" + LIBDIR, "-lmytrim", "-L" + GPULIB, "-lmytrim_gpu",
                "-Wl,-rpath," + LIBDIR, "-Wl,-rpath," + GPULIB]
    else:
        cmd += ["-c", src, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return os.path.basename(src), r.returncode, r.stderr[-1500:]


def test_reference_callers_found():
    assert len(DRIVERS) >= 12 and len(MODULES) >= 3


@pytest.mark.parametrize("kind", ["drivers", "modules"])
def test_reference_callers_build_unchanged(tmp_path, kind):
    assert os.path.exists(os.path.join(LIBDIR, "libmytrim.so")), "build mytrim_amd/host first"
    srcs = DRIVERS if kind == "drivers" else MODULES
    jobs = [(s, str(tmp_path / os.path.basename(s)[:-2]), kind == "drivers") for s in srcs]
    with ThreadPoolExecutor(min(8, os.cpu_count() or 1)) as ex:
        res = list(ex.map(lambda j: _compile(*j), jobs))
    bad = [(n, e) for n, rc, e in res if rc != 0]
    assert not bad, bad
    if kind == "drivers":
        for s, out, _ in jobs:
            assert os.access(out, os.X_OK), out
