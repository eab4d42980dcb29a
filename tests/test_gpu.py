"""GPU parity tests: the HIP library (through its C ABI) against the oracle,
the compiled reference's fixtures and the reference's statistics.

Tolerances (FP64 path):
* stopping (getrstop): relative 1e-12 against the reference's own values.
* scattering: |d cos(theta)| <= 1e-12 and |d s2| <= 1e-12 s2 + 8 ulp(1); s2 =
  1 - co^2 is formed by a subtraction from 1, so glancing collisions carry an
  absolute rounding floor whatever the libm; sin(theta) to 1e-12 where it is
  >= 1e-2, below that through sin^2 = 1 - cos^2.
* cascades: the device draw scheme is restated by the oracle (rng=1), so
  histories are compared one by one; the device's elementary functions and
  the host libm differ in the last ulps, which can flip a rare threshold
  decision deep in a long cascade, so
  >= 97 % of histories must agree exactly in every integer tally and totals
  within 1 %.
* against the reference (std::mt19937 streams): the two-sample criteria of
  tally_stats (test_gpu_stats.py); the vacancy table within 3 sigma + 1.5 %.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle_lib as O
from fixture_cases import ENERGIES, IONS, TARGETS, c4_uniform, trace_cases
from mytrim_amd import configs, gpu
from mytrim_amd.problem import (CUT, GEO_SOLID, PBC, Bins, Element, Geometry, Material,
                                Primaries, Problem)

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def eng():
    e = gpu.Engine(0)
    yield e
    e.close()


def load_tally(name):
    with open(os.path.join(GOLD, "tally_%s.json" % name)) as f:
        return json.load(f)


def target_problem(rho, els):
    return Problem(materials=[Material(rho, [Element(Z, m, t) for Z, m, t in els])],
                   geometry=Geometry(), primaries=Primaries.single(1, 1.0, 1.0, (0, 0, 0)),
                   n_hist=0)


# ---------------------------------------------------------------- per-function twins
def test_stopping_matches_reference_grid(eng):
    g = np.load(os.path.join(GOLD, "stopping_grid.npz"))
    worst = 0.0
    for j, (_, rho, els) in enumerate(TARGETS):
        eng.set_target(target_problem(rho, els))
        for i, (Z1, m1) in enumerate(IONS):
            se = eng.stopping(Z1, m1, 0, ENERGIES)
            ref = g["se"][i, j]
            worst = max(worst, float(np.max(np.abs(se - ref) / np.abs(ref))))
    print("stopping worst relative error %.3g" % worst)
    assert worst <= 1e-12


@pytest.mark.parametrize("potential", [0, 1, 2])
def test_scatter_matches_oracle(eng, potential):
    eps = np.repeat(np.logspace(-5, 3, 81), 41)
    b = np.tile(np.concatenate([[1e-6, 1e-3], np.linspace(0.01, 20.0, 39)]), 81)
    s2, ct, st, it = eng.scatter(eps, b, potential)
    ref = np.array([O.scatter(e, bb, potential) for e, bb in zip(eps, b)])
    d_ct = np.abs(ct - ref[:, 1])
    d_s2 = np.abs(s2 - ref[:, 0])
    print("max |d ct| %.3g, max |d s2| - tol %.3g" % (d_ct.max(), (d_s2 - 1e-12 * ref[:, 0]).max()))
    assert np.all(d_ct <= 1e-12)
    assert np.all(d_s2 <= 1e-12 * ref[:, 0] + 8 * 2.0 ** -52)
    # st = sqrt(1 - ct^2) (trim.C:178, 271) amplifies a difference in ct by ct/st:
    # where st >= 1e-2 it must match to 1e-12; everywhere st^2 = 1 - ct^2 must
    # differ by no more than the ct difference explains, 2 |d ct| + rounding
    good = ref[:, 2] >= 1e-2
    assert np.all(np.abs(st - ref[:, 2])[good] <= 1e-12 + 1e-12 * ref[:, 2][good])
    assert np.all(np.abs(st * st - ref[:, 2] ** 2) <= 2 * d_ct + 8 * 2.0 ** -52)
    assert np.mean(it == ref[:, 3]) >= 0.999


# ---------------------------------------------------------------- single precision
# The single-precision device functions of the FP32 cascade mode against the
# reference's FP64 values (FP32 unit roundoff 6e-8).
# * stopping: about thirty operations, the powers as exp2(y log2 x), whose
#   rounding grows with |y log x| (up to ~10 in the effective-charge fits):
#   relative 2e-5.
# * scattering: single precision cannot resolve s2 = 1 - co^2 for glancing
#   collisions, so the FP32 path forms s2 from the apsis identity
#   r - b = v r^2 / (eps (r + b)), exact at the apsis.  The reference
#   evaluates its formula at the Newton iterate that stopped at a relative
#   step of 1e-3 (trim.C:233), not at the apsis: in FP64 the two formulas
#   differ by up to 5.1e-4 in s2 for the universal and Moliere potentials,
#   and single precision adds no more than that: relative 1e-3 in s2, 1e-4
#   in cos theta.
#   The reference's C-Kr derivative coefficients are not the derivative of its
#   screening (trim.C:215-221: 0.531865 for 0.191 x 0.279 = 0.0533, 0.6437
#   for 0.335 x 1.1919 = 0.3993), so its iteration converges only linearly
#   and stops short of the apsis by up to 6e-2 in s2: bound 8e-2 there.
FP32_STOP_RTOL = 2e-5
FP32_S2_RTOL = {0: 1e-3, 1: 1e-3, 2: 8e-2}
FP32_CT_ATOL = {0: 1e-4, 1: 1e-4, 2: 8e-2}


def test_fp32_stopping_within_tolerance_of_reference(eng):
    g = np.load(os.path.join(GOLD, "stopping_grid.npz"))
    errs = []
    for j, (_, rho, els) in enumerate(TARGETS):
        eng.set_target(target_problem(rho, els))
        for i, (Z1, m1) in enumerate(IONS):
            se = eng.stopping(Z1, m1, 0, ENERGIES, arith=gpu.MT_FP32)
            ref = g["se"][i, j]
            errs.append(np.abs(se - ref) / np.abs(ref))
    e = np.concatenate(errs)
    print("fp32 stopping relative error: median %.2e p99 %.2e max %.2e" % (
        np.median(e), np.quantile(e, 0.99), e.max()))
    assert e.max() <= FP32_STOP_RTOL


@pytest.mark.parametrize("potential", [0, 1, 2])
def test_fp32_scatter_within_tolerance_of_oracle(eng, potential):
    eps = np.repeat(np.logspace(-5, 3, 81), 41)
    b = np.tile(np.concatenate([[1e-6, 1e-3], np.linspace(0.01, 20.0, 39)]), 81)
    # the inputs the FP32 functions see (rounded to single precision)
    eps32, b32 = eps.astype(np.float32).astype(float), b.astype(np.float32).astype(float)
    s2, ct, st, it = eng.scatter(eps, b, potential, arith=gpu.MT_FP32)
    ref = np.array([O.scatter(e, bb, potential) for e, bb in zip(eps32, b32)])
    rs2 = np.abs(s2 - ref[:, 0]) / np.maximum(ref[:, 0], 1e-300)
    dct = np.abs(ct - ref[:, 1])
    print("fp32 scatter (potential %d): |d s2|/s2 median %.2e p99 %.2e max %.2e; max |d ct| %.2e" % (
        potential, np.median(rs2), np.quantile(rs2, 0.99), rs2.max(), dct.max()))
    assert rs2.max() <= FP32_S2_RTOL[potential]
    assert dct.max() <= FP32_CT_ATOL[potential]
    assert np.all(np.abs(st * st + ct * ct - 1.0) <= 1e-5)


# ---------------------------------------------------------------- cascades vs oracle
CASES = {
    "c1": lambda: configs.c1_he_fe(400, seed=5),
    "c2": lambda: configs.c2_cu_cu(6, seed=5),
    "c3": lambda: configs.c3_ag_layers(8, seed=5),
    "c4": lambda: c4_uniform(2000.0, 12).with_(seed=5),
    "c4mix": lambda: configs.c4_mixed_uo2(64, seed=5, w=2000.0),
    "c5": lambda: configs.c5_c_w(24, seed=5),
    "kp": lambda: configs.c2_cu_cu(200, seed=5).with_(follow_max_gen=1, kp=True),
    "moliere": lambda: configs.c1_he_fe(100, seed=5).with_(potential=1),
    "ckr": lambda: configs.c5_c_w(10, seed=5).with_(potential=2),
    "wire": lambda: configs.wire_b_si(400, seed=5),
    "bwire": lambda: configs.burried_wire_p_si(24, seed=5),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_cascade_histories_match_oracle(eng, case):
    prob = CASES[case]()
    g = eng.run(prob, per_history=True)
    o = O.run(prob, rng=1, nthreads=8, per_history=True)
    keys = ["h_vac", "h_coll", "h_flights", "h_state"]
    same = np.ones(prob.n_hist, bool)
    for k in keys:
        same &= g[k] == o[k]
    pos_ok = np.isclose(g["h_x"], o["h_x"], rtol=1e-9, atol=1e-6)
    frac = float(np.mean(same & pos_ok))
    print("%s: %d histories, %d collisions, %.1f ms, identical fraction %.4f"
          % (case, prob.n_hist, g["n_coll"], g["kernel_ms"], frac))
    assert frac >= 0.97
    for k in ["n_coll", "n_flight", "vac", "disp", "range_n"]:
        assert g[k] == pytest.approx(o[k], rel=0.01), k
    for k in ["Eel", "Enuc"]:
        assert g[k] == pytest.approx(o[k], rel=0.01), k
    tot = g["Eel"] + g["Enuc"] + g["Eunf"] + g["Elost"] + g["Edrop"] + g["Evac"]
    assert tot == pytest.approx(g["E0"], rel=1e-11)


@pytest.mark.parametrize("case", ["c1", "c4mix", "c5"])
def test_wave_roles_leave_results_unchanged(eng, case):
    """Feeder waves (primaries only, every recoil handed to the HBM queue)
    change which wave runs a flight, not what the flight computes: integer
    tallies and histograms equal those without roles, energies to 1e-12."""
    prob = CASES[case]()
    eng.tune()
    a = eng.run(prob)
    try:
        for k64 in (8, 32):
            eng.tune(primary_waves64=k64)
            b = eng.run(prob)
            print("%s roles %d/64: %.2f ms (without roles %.2f ms), primaries' share of collisions %.3f" % (
                case, k64, b["kernel_ms"], a["kernel_ms"], b["coll_primary"] / b["n_coll"]))
            for k in ["n_coll", "n_flight", "vac", "repl", "subst", "inter", "lost", "vacuum", "disp", "range_n",
                      "coll_primary"]:
                assert a[k] == b[k], k
            for k in ["range_hist", "lat_y_hist", "lat_z_hist", "vac_hist", "repl_hist"]:
                assert np.array_equal(a[k], b[k]), k
            for k in ["Eel", "Enuc", "Eunf", "Elost", "Edrop", "Evac"]:
                assert b[k] == pytest.approx(a[k], rel=1e-12, abs=1e-9), k
    finally:
        eng.tune()


def test_cascade_bitwise_totals_when_histories_agree(eng):
    """Where every history agrees, the tallies must agree exactly."""
    prob = configs.c1_he_fe(300, seed=9)
    g = eng.run(prob, per_history=True)
    o = O.run(prob, rng=1, nthreads=8, per_history=True)
    if not np.array_equal(g["h_coll"], o["h_coll"]) or not np.array_equal(g["h_vac"], o["h_vac"]):
        pytest.skip("a history diverged at the last ulp; covered by the fraction test")
    for k in ["n_coll", "n_flight", "vac", "repl", "subst", "inter", "lost", "disp", "range_n"]:
        assert g[k] == o[k], k
    for k in ["range_hist", "lat_y_hist", "vac_hist", "repl_hist"]:
        assert np.array_equal(g[k], o[k]), k
    assert g["Eel"] == pytest.approx(o["Eel"], rel=1e-12)


def test_history_offsets_select_the_same_streams(eng):
    """History h's result does not depend on the batch it is run in."""
    full = configs.c1_he_fe(200, seed=21)
    a = eng.run(full, per_history=True)
    b = eng.run(full.with_(hist0=100, n_hist=100), per_history=True)
    assert np.array_equal(a["h_vac"][100:], b["h_vac"])
    assert np.array_equal(a["h_coll"][100:], b["h_coll"])


# ---------------------------------------------------------------- vs the reference
# (the two-sample tally comparisons are in test_gpu_stats.py)
def test_vacancy_table_on_gpu(eng):
    rows = np.loadtxt(os.path.join(GOLD, "validation", "vacancy_count_comparison.dat"),
                      skiprows=2, delimiter=",")
    for (ion, e, val) in [("cu", 1e5, rows[4, 10]), ("si", 1e5, rows[4, 2]), ("xe", 1e5, rows[4, 6])]:
        n = 2000
        g = eng.run(configs.vacancy_case(ion, e, n, seed=3), per_history=True)
        mean = g["vac"] / n
        sem = np.std(g["h_vac"]) / math.sqrt(n)
        assert abs(mean - val) < 3 * sem + 0.015 * val, (ion, e, mean, val)


# ---------------------------------------------------------------- drop-in flight path
def replay_history(eng, prob, h, max_rows):
    """TrimBase::trim one flight at a time on the GPU, FIFO drain, with the
    reference's std::mt19937 stream; returns trace rows like orc_trace."""
    seed = O.mt_history_seed(prob.seed, h)
    stream = O.mt_stream(seed, 400000)
    pos = 0
    p = prob.primaries
    q = [dict(Z=int(p.Z[0]), m=float(p.m[0]), E=float(p.E[0]), pos=list(p.pos[0]),
              dir=list(p.dir[0]), gen=0, tag=-1)]
    rows = []
    flight = 0
    els = [e for mat in prob.materials for e in mat.elements]
    offs = np.cumsum([0] + [len(m.elements) for m in prob.materials])
    while q and len(rows) < max_rows:
        d = q.pop(0)
        ion = gpu.Ion(d["Z"], d["m"], d["E"], tuple(d["pos"]), tuple(d["dir"]), d["gen"], d["tag"],
                      0, prob.Ef)
        used, evs = eng.trim_flight(ion, stream[pos:pos + 200000])
        pos += used
        for ic, e in enumerate(evs, 1):
            el = els[offs[e.material] + e.element]
            rows.append([flight, ic, d["Z"], e.pka_E, *e.pka_pos, *e.pka_dir, e.pka_state, e.ls,
                         e.dee, e.den, el.Z, e.rec_E, *e.rec_dir, d["gen"], 0.0])
            if e.fate == 1:
                rd = np.array(e.rec_dir)
                inv = 1.0 / math.sqrt(rd[0] * rd[0] + rd[1] * rd[1] + rd[2] * rd[2])
                q.append(dict(Z=el.Z, m=el.m, E=e.rec_E, pos=list(e.rec_pos),
                              dir=list(rd * inv), gen=d["gen"] + 1, tag=e.mat_tag))
        flight += 1
    return np.array(rows[:max_rows])


@pytest.mark.parametrize("case", ["c1", "c2", "c3", "c4", "c5", "moliere", "ckr", "wire", "bwire"])
def test_flight_path_reproduces_reference_trace(eng, case):
    ref = np.load(os.path.join(GOLD, "trace_%s_h0.npz" % case))["rows"][:600]
    prob = trace_cases()[case]
    eng.set_target(prob)
    mine = replay_history(eng, prob, 0, len(ref))
    assert mine.shape == ref.shape
    # integer columns exact: flight, ic, Z, state, element Z, gen
    for c in (0, 1, 2, 10, 14, 19):
        assert np.array_equal(mine[:, c], ref[:, c]), c
    scale = np.maximum(np.abs(ref), 1.0)
    err = np.max(np.abs(mine - ref) / scale, axis=0)
    print("%s: max scaled deviation per column %s" % (case, np.array2string(err, precision=1)))
    assert np.all(err <= 1e-9)


# ---------------------------------------------------------------- edge cases
def test_zero_histories(eng):
    g = eng.run(configs.c1_he_fe(0))
    assert g["n_coll"] == 0 and g["n_flight"] == 0


def test_ions_leaving_through_the_cut_face(eng):
    prob = configs.c1_he_fe(64)
    prob = prob.with_(primaries=Primaries.single(2, 4.0026, 1e5, (0.0, 5000.0, 5000.0), (-1, 0, 0)))
    g = eng.run(prob, per_history=True)
    o = O.run(prob, rng=1, per_history=True)
    assert g["lost"] == o["lost"] == 64
    assert np.array_equal(g["h_coll"], o["h_coll"])
    assert g["range_n"] == 0


def test_sub_threshold_primaries(eng):
    prob = configs.c2_cu_cu(128).with_(
        primaries=Primaries.single(29, 63.546, 2.0, (100.0, 1e4, 1e4)))
    g = eng.run(prob, per_history=True)
    o = O.run(prob, rng=1, per_history=True)
    assert np.array_equal(g["h_coll"], o["h_coll"])
    assert g["inter"] == o["inter"]


def test_non_finite_energy_is_reported():
    """A NaN stopping coefficient makes every flight's energy NaN: the run
    returns MT_E_DEVICE "non-finite-energy" at once (not a runaway after 1e7
    collisions per flight), every kernel exits, and the context runs
    normally again once the tables are restored.  The single-flight path
    reports the same."""
    rows = gpu.default_tables()
    bad = rows.copy()
    bad[28, 7:15] = np.nan  # Cu (Z = 29): the eight stopping coefficients
    e = gpu.Engine(0)
    try:
        prob = configs.c2_cu_cu(256, seed=5)
        e.set_tables(bad)
        e.set_target(prob)
        with pytest.raises(gpu.MtError, match="non-finite-energy"):
            e.run(prob)
        ion = gpu.Ion(29, 63.546, 5.0e4, (0.0, 100.0, 100.0), (1.0, 0.0, 0.0), 0, -1, 0, 3.0)
        with pytest.raises(gpu.MtError, match="non-finite-energy"):
            e.trim_flight(ion, np.full(4096, 0.5))
        e.set_tables(rows)
        e.set_target(prob)
        r = e.run(prob)
        ref = eng_reference_run(prob)
        assert r["n_coll"] == ref["n_coll"] and r["vac"] == ref["vac"]
    finally:
        e.close()


def eng_reference_run(prob):
    e = gpu.Engine(0)
    try:
        return e.run(prob)
    finally:
        e.close()


def test_invalid_input_raises(eng):
    bad = configs.c1_he_fe(4).with_(materials=[Material(7.87, [Element(120, 55.0)])])
    with pytest.raises(gpu.MtError):
        eng.set_target(bad)
    for kw in ({"block": 512}, {"block": 100}, {"queue_cap": 1000}, {"stack_cap": 5000}):
        with pytest.raises(gpu.MtError):
            eng.tune(**kw)
    eng.tune()
    # primaries the kernel cannot fly are refused before any launch
    base = configs.c1_he_fe(4)
    for E, pos, d, msg in ((float("inf"), (0, 5000, 5000), (1, 0, 0), "energy"),
                           (-1.0, (0, 5000, 5000), (1, 0, 0), "energy"),
                           (1e5, (0, float("nan"), 5000), (1, 0, 0), "position"),
                           (1e5, (0, 5000, 5000), (0, 0, 0), "direction"),
                           (1e5, (0, 5000, 5000), (float("inf"), 0, 0), "direction")):
        with pytest.raises(gpu.MtError, match=msg):
            eng.run(base.with_(primaries=Primaries.single(2, 4.0026, E, pos, d)))
    assert eng.run(base)["n_coll"] > 0


# ---------------------------------------------------------------- FP32 physics path
@pytest.mark.parametrize("name,prob", [
    ("c1", lambda: configs.c1_he_fe(20000, seed=41)),
    ("c3", lambda: configs.c3_ag_layers(300, seed=41)),
    ("c5", lambda: configs.c5_c_w(1000, seed=41)),
])
def test_fp32_tallies_within_one_percent_of_fp64(eng, name, prob):
    """FP32 stopping and scattering, FP64 energies, positions and ledger:
    vacancies/ion, mean range and the electronic share of the energy within
    1 % and within 2 sigma of the FP64 path on the same histories."""
    p = prob()
    a = eng.run(p, per_history=True, arith=gpu.MT_FP64)
    b = eng.run(p, per_history=True, arith=gpu.MT_FP32)
    n = p.n_hist
    va, vb = a["vac"] / n, b["vac"] / n
    sig = math.sqrt((np.var(a["h_vac"]) + np.var(b["h_vac"])) / n)
    ra = a["range_sum"] / a["range_n"]
    rb = b["range_sum"] / b["range_n"]
    rs = math.sqrt((a["range_sum2"] / a["range_n"] - ra ** 2) / a["range_n"] * 2)
    print("%s FP64 %.1f ms FP32 %.1f ms; vac/ion %.3f vs %.3f; range %.2f vs %.2f; Eel %.5g vs %.5g"
          % (name, a["kernel_ms"], b["kernel_ms"], va, vb, ra, rb, a["Eel"], b["Eel"]))
    assert abs(vb - va) <= 0.01 * va and abs(vb - va) <= 2 * sig
    assert abs(rb - ra) <= 0.01 * abs(ra) and abs(rb - ra) <= 2 * rs
    assert b["Eel"] / b["E0"] == pytest.approx(a["Eel"] / a["E0"], rel=0.01)
    tot = b["Eel"] + b["Enuc"] + b["Eunf"] + b["Elost"] + b["Edrop"] + b["Evac"]
    assert tot == pytest.approx(b["E0"], rel=1e-11)


# ---------------------------------------------------------------- launch machinery
def test_graph_replays_match_direct_launches(eng):
    """The cached run graph (reset kernel, persistent kernel, flush kernel,
    copies to pinned memory), replayed several times with direct launches and
    host work in between, gives the direct launch's tallies every time."""
    prob = configs.c1_he_fe(400, seed=77)
    eng.tune(graph=-1)
    ref = eng.run(prob, per_history=True)
    try:
        for rep in range(3):
            eng.tune(graph=0)
            g = eng.run(prob, per_history=True)
            eng.tune(graph=-1)
            d = eng.run(prob.with_(n_hist=50, hist0=1000 + rep), per_history=True)   # eager work between
            assert d["n_coll"] > 0
            for k in ["n_coll", "n_flight", "vac", "repl", "subst", "inter", "lost", "disp", "range_n"]:
                assert g[k] == ref[k], (rep, k)
            for k in ["range_hist", "lat_y_hist", "lat_z_hist", "vac_hist", "repl_hist", "h_vac", "h_coll",
                      "h_flights", "h_state"]:
                assert np.array_equal(g[k], ref[k]), (rep, k)
            for k in ["Eel", "Enuc", "Eunf", "Elost", "Edrop", "Evac", "range_sum"]:
                assert g[k] == pytest.approx(ref[k], rel=1e-12, abs=1e-9), (rep, k)
    finally:
        eng.tune()


def test_graph_survives_shape_changes(eng):
    """Histogram sizes and batch sizes change between calls: the graph is
    rebuilt or updated and the tallies stay those of a direct launch."""
    a = configs.c1_he_fe(120, seed=5)
    b = configs.c5_c_w(30, seed=5)
    out = {}
    try:
        for mode in (-1, 0):
            eng.tune(graph=mode)
            out[mode] = [eng.run(p, per_history=True) for p in (a, b, a.with_(n_hist=60))]
    finally:
        eng.tune()
    for x, y in zip(out[-1], out[0]):
        assert x["n_coll"] == y["n_coll"] and x["vac"] == y["vac"]
        assert np.array_equal(x["range_hist"], y["range_hist"])
        assert np.array_equal(x["h_coll"], y["h_coll"])


def test_autotune_picks_a_working_shape(eng):
    prob = configs.c1_he_fe(3000, seed=3)
    best, rate = eng.autotune(prob, sample=500)
    print("autotune: block %d, blocks/CU %d, stack %d -> %.3g coll/s" %
          (best.block, best.blocks_per_cu, best.stack_cap, rate))
    assert best.block in (64, 128, 256) and best.stack_cap in (64, 96, 128, 160, 192, 256, 384)
    assert rate > 0
    r = eng.run(prob.with_(n_hist=300), per_history=True)
    eng.tune()
    s = eng.run(prob.with_(n_hist=300), per_history=True)
    assert np.array_equal(r["h_coll"], s["h_coll"]) and r["vac"] == s["vac"]
    # the launch record's mean shader clock (MI355X: up to 2.4 GHz)
    assert 300.0 < r["sclk_mhz"] < 3000.0 and 300.0 < s["sclk_mhz"] < 3000.0


def test_depth_bins_split_between_lds_and_hbm(eng):
    """Vacancy/replacement depth bins: all in HBM, the first few in LDS and the
    rest in HBM, or as many in LDS as fit (three blocks per CU asked for):
    identical histograms."""
    prob = configs.c2_cu_cu(200, seed=5)
    out = []
    try:
        for lh, bpc in ((-1, 0), (7, 0), (300, 0), (0, 3), (0, 0)):
            eng.tune(block=256, blocks_per_cu=bpc, stack_cap=128, lds_hist=lh)
            out.append(eng.run(prob))
    finally:
        eng.tune()
    assert out[0]["vac_hist"].sum() > 0 and out[0]["repl_hist"].sum() > 0
    for r in out[1:]:
        assert np.array_equal(r["vac_hist"], out[0]["vac_hist"])
        assert np.array_equal(r["repl_hist"], out[0]["repl_hist"])
        assert r["vac"] == out[0]["vac"] and r["repl"] == out[0]["repl"]


@pytest.mark.parametrize("block,cap", [(64, 256), (64, 512), (128, 256)])
def test_deep_lds_stacks_match_oracle(eng, block, cap):
    """LDS stacks deeper than a wave: refills still move at most one entry per
    lane, and the histories stay those of the oracle."""
    prob = configs.c2_cu_cu(4, seed=8)
    eng.tune(block=block, stack_cap=cap)
    try:
        g = eng.run(prob, per_history=True)
    finally:
        eng.tune()
    o = O.run(prob, rng=1, nthreads=4, per_history=True)
    assert np.array_equal(g["h_vac"], o["h_vac"]) and np.array_equal(g["h_coll"], o["h_coll"])


# ---------------------------------------------------------------- event streams
@pytest.mark.parametrize("case", ["vac_cu_10keV", "wire", "c4"])
def test_event_streams_match_oracle(eng, case):
    """Device records of TrimDefectLog / TrimHistory / TrimPhononOut and the
    TrimVacMap map against the oracle on the device draw scheme: per history
    the same multiset of records (ids aside: the device numbers flights by
    their key, the reference by a global counter)."""
    from fixture_cases import log_cases
    prob, vz = log_cases()[case]
    g = eng.run(prob, log=dict(mask=15, max_records=1 << 20, vacmap_Z=vz))
    rec, vm = O.run_log(prob, 7, vz, rng=1)
    d = g["records"]
    assert g["n_records"] == len(d) == len(rec)
    same = 0
    for h in range(prob.hist0, prob.hist0 + prob.n_hist):
        a, b = d[d["hist"] == h], rec[rec["hist"] == h]
        if len(a) != len(b):
            continue

        def key(r):
            return np.lexsort((r["E"], r["pos"][:, 2], r["pos"][:, 1], r["pos"][:, 0], r["gen"], r["Z"], r["kind"]))
        a, b = a[key(a)], b[key(b)]
        ints_ok = all(np.array_equal(a[k], b[k]) for k in ("kind", "Z", "gen", "tag"))
        fl = np.column_stack([a["pos"], a["E"], a["value"], a["m"]])
        fr = np.column_stack([b["pos"], b["E"], b["value"], b["m"]])
        if ints_ok and np.all(np.abs(fl - fr) <= 1e-9 * np.maximum(np.abs(fr), 1.0)):
            same += 1
    print("%s: %d records, %d of %d histories identical" % (case, len(d), same, prob.n_hist))
    assert same >= 0.97 * prob.n_hist
    if same == prob.n_hist:
        assert np.array_equal(g["vacmap"], vm)
    assert g["vacmap"].sum() > 0
    # the ion at rest: one I/S/R record per resting flight, matching the tallies
    assert np.sum(np.isin(d["kind"], (2, 3, 4))) == g["repl"] + g["subst"] + g["inter"]
    assert np.sum(d["kind"] == 1) == g["vac"]


@pytest.mark.parametrize("case", ["c3", "xe_si_sio2"])
def test_layer_transport_matches_oracle(eng, case):
    """SampleDynamic's atom moves per (layer, Z) (MT_LOG_TRANSPORT) against the
    oracle on the device draw scheme: equal when every history ran the same
    flights (per-history collision and vacancy counts), within the moves of
    the histories that diverged otherwise."""
    from fixture_cases import transport_cases
    prob = transport_cases()[case]
    g = eng.run(prob, per_history=True, log=dict(mask=16))
    _, _, tr = O.run_log(prob, 0, rng=1, transport=True)
    o = O.run(prob, rng=1, nthreads=4, per_history=True)
    same = (g["h_coll"] == o["h_coll"]) & (g["h_vac"] == o["h_vac"])
    print("%s: %d of %d histories identical, %d moves" % (case, same.sum(), prob.n_hist, np.abs(tr).sum()))
    assert same.mean() >= 0.95
    assert g["transport"].shape == tr.shape
    if same.all():
        assert np.array_equal(g["transport"], tr)
    else:
        assert np.abs(g["transport"] - tr).sum() <= 0.05 * np.abs(tr).sum()
    # the recoil flights leave their layers: one -1 per flight with gen > 0,
    # none for the primaries
    assert (tr < 0).any() and (tr > 0).any()
    # no other tally changes with the transport tally on
    plain = eng.run(prob)
    for k in ("n_coll", "n_flight", "vac", "repl", "inter"):
        assert plain[k] == g[k], k


def test_layer_transport_needs_layers(eng):
    prob = configs.c2_cu_cu(2)  # SampleSolid
    with pytest.raises(RuntimeError, match="layers"):
        eng.run(prob, log=dict(mask=16))


def test_event_stream_capacity_and_counters(eng):
    """A too-small record buffer keeps the first records and reports the total."""
    prob = configs.vacancy_case("cu", 1e4, 8)
    full = eng.run(prob, log=dict(mask=7, max_records=1 << 18))
    part = eng.run(prob, log=dict(mask=7, max_records=100))
    assert part["n_records"] == full["n_records"] > 100
    assert len(part["records"]) == 100
    plain = eng.run(prob)
    for k in ("n_coll", "vac", "repl", "inter"):
        assert plain[k] == full[k], k
