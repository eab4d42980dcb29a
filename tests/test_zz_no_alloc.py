"""mt_run_cascades_noalloc (device memory from the device's static scratch,
no allocation) against mt_run_cascades on the same histories, through the C
ABI: two calls in a row, the second smaller than the first (the scratch
keeps the first call's contents, so this checks that everything that must
start at zero is cleared).  Flights draw from counter-based streams keyed by
history, so the integer tallies and histograms are equal exactly; the energy
sums are compensated sums merged in another order (1e-12 relative).  Also the
limits: more histories or per-history primaries than the scratch holds are
refused with MT_E_LIMIT, per-history records with MT_E_ARG."""
import ctypes as C

import numpy as np
import pytest

from mytrim_amd import configs, gpu

pytestmark = pytest.mark.gpu

INT_KEYS = ["n_coll", "n_flight", "vac", "repl", "subst", "inter", "lost", "vacuum", "rutherford",
            "newton", "disp", "range_n"]
HISTS = ["range_hist", "lat_y_hist", "lat_z_hist", "vac_hist", "repl_hist"]


@pytest.fixture(scope="module")
def eng():
    e = gpu.Engine(0)
    yield e
    e.close()


def same(a, b):
    for k in INT_KEYS:
        assert a[k] == b[k], k
    for k in HISTS:
        assert np.array_equal(a[k], b[k]), k
    for k in ("Eel", "Enuc", "Eunf", "Elost", "Edrop", "Evac", "range_sum", "E0"):
        assert a[k] == pytest.approx(b[k], rel=1e-12, abs=1e-9), k


@pytest.mark.parametrize("name", ["c2", "c4mix"])
def test_noalloc_matches_run_cascades_twice(eng, name):
    if name == "c2":
        big, small = configs.c2_cu_cu(4000, seed=19), configs.c2_cu_cu(700, seed=23)
    else:
        p = configs.c4_mixed_uo2(6000, seed=19, w=4000.0)
        big, small = p, p.batch(1000, 900)
    for prob in (big, small):
        a = eng.run(prob, noalloc=True)
        b = eng.run(prob)
        print("%s %d histories: %d collisions, noalloc %.2f ms, run %.2f ms, queue peak %d" % (
            name, prob.n_hist, a["n_coll"], a["kernel_ms"], b["kernel_ms"], a["queue_peak"]))
        same(a, b)


def test_noalloc_limits(eng):
    prob = configs.c1_he_fe(1 << 22, seed=3)
    eng.set_target(prob)
    with pytest.raises(gpu.MtError, match="2\\^22"):
        eng.run(prob.with_(n_hist=(1 << 22) + 1), noalloc=True)
    mixed = configs.c4_mixed_uo2((1 << 18) + 1, seed=3, w=2000.0)
    with pytest.raises(gpu.MtError, match="2\\^18"):
        eng.run(mixed, noalloc=True)
    with pytest.raises(ValueError):
        eng.run(configs.c1_he_fe(10, seed=3), noalloc=True, per_history=True)
    # the entry point itself refuses per-history records
    small = configs.c1_he_fe(10, seed=3)
    pk, o, keep = eng._launch_args(small, True, gpu.MT_FP64)
    out = gpu.Tallies()
    rc = gpu.lib().mt_run_cascades_noalloc(eng.h, C.byref(pk), C.byref(o), C.byref(out))
    assert rc == -1
    # and the engine still runs normally afterwards
    r = eng.run(small, noalloc=True)
    assert r["n_coll"] == eng.run(small)["n_coll"]
