"""The C++ API built as inside MOOSE/Magpie (MYTRIM_ENABLED).

The reference switches its types and error reporting at compile time
(simconf.h:31-42): Real, Point and Utility::pow come from MOOSE and libMesh,
and invalid stoichiometry or unreadable tables reach mooseError
(material.C:50-71, simconf.C:142-165) instead of printing and exiting.  MOOSE
and libMesh are not installed here, so tests/native/moose_double holds test
doubles of the few declarations involved.  This test compiles every source of
the C++ API and the build's own drivers in that mode, links a driver that runs
the error paths needing no GPU (tests/native/moose_mode.cpp), and checks that
mooseError received the reference's messages.  A build inside a real MOOSE
tree is not tested."""
import glob
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "mytrim_amd", "host")
GPULIB = os.path.join(ROOT, "mytrim_amd", "_build")
DOUBLE = os.path.join(ROOT, "tests", "native", "moose_double")
FLAGS = ["g++", "-std=c++17", "-O1", "-Wall", "-DMYTRIM_ENABLED", "-I" + DOUBLE,
         "-I" + os.path.join(HOST, "include"), "-I" + os.path.join(HOST, "src"), "-I" + os.path.join(ROOT, "include")]


def _obj(src, tmp):
    out = os.path.join(tmp, os.path.basename(src) + ".o")
    r = subprocess.run(FLAGS + ["-c", src, "-o", out], capture_output=True, text=True)
    return src, out, r.returncode, r.stderr[-1500:]


def test_api_builds_and_reports_through_moose_error(tmp_path):
    assert os.path.exists(os.path.join(GPULIB, "libmytrim_gpu.so")), "build mytrim_amd first"
    srcs = sorted(glob.glob(os.path.join(HOST, "src", "*.cpp")))
    drivers = [os.path.join(ROOT, "tests", "native", n) for n in ("moose_mode.cpp", "fission_bubbles.cpp")]
    with ThreadPoolExecutor(min(8, os.cpu_count() or 1)) as ex:
        res = list(ex.map(lambda s: _obj(s, str(tmp_path)), srcs + drivers))
    bad = [(os.path.basename(s), e) for s, _, rc, e in res if rc != 0]
    assert not bad, bad
    lib_objs = [o for s, o, _, _ in res if s in srcs]
    exe = str(tmp_path / "moose_mode")
    r = subprocess.run(FLAGS + ["-o", exe, res[len(srcs)][1]] + lib_objs +
                       ["-L" + GPULIB, "-lmytrim_gpu", "-Wl,-rpath," + GPULIB], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    env = {k: v for k, v in os.environ.items() if k != "MYTRIM_DATADIR"}
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60, env=env)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.splitlines()
    assert lines[0].startswith("libMesh point norm 3.0")
    assert lines[1] == "mooseError: Stoichiometry invalid, all elements zero."
    assert lines[2] == "mooseError: Error reading /nonexistent-mytrim-data/SCOEF.95A"
