#!/usr/bin/env python3
"""Write the golden fixtures from the reference compiled from its own sources.

Runs only where the reference is mounted (oracle/Makefile target `ref` builds
oracle/_ref/ref_harness from /root/reference).  Every number stored here is
computed by the reference library; the harness only drives its API.

  stopping_grid.npz   MaterialBase::getrstop (material.C:114-122) on an
                      (ion, target, energy) grid
  average.npz         MaterialBase::prepare/average constants (material.C:37-110)
  trace_<case>.npz    per-collision state of TrimBase::trim (trim.C:35-425) for
                      the first collisions of one history, std::mt19937 stream
  tally_<case>.json   full-cascade tallies (FIFO drain of apps/runmytrim.C:71-93),
                      one std::mt19937 seed per history
  log_<case>.npz      what TrimDefectLog and TrimPhononOut write (17 digits), the
                      positions TrimHistory stores and TrimVacMap's map
                      (trim.C:466-527, trim.h:139-227), histories in order
  stat_<case>.npz     larger full-cascade samples for the two-sample tests:
                      per-history vacancies, collisions, Eel, Enuc and E0,
                      the end-point histograms (range, lateral) and the
                      depth histograms (vacancies, replacements) of each of
                      STAT_BATCHES equal batches of histories
  runmytrim/<case>/   a JSON deck of the reference (tests/json, validation/) with
                      fewer ions, and the output file and totals the reference's
                      runmytrim loop writes for it (oracle/_ref/runmytrim_ref:
                      the reference's driver loop and output modules; jsoncpp is
                      absent, so the deck is handed over as a flat spec)
Usage:  python tests/golden/make_fixtures.py [--quick] [--tally=NAME] [--trace=NAME] [--logs] [--runmytrim]
                                            [--stat[=NAME]]
"""
import json
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle_lib as O  # noqa: E402

HARNESS = O.REF_HARNESS

from fixture_cases import (ENERGIES, IONS, STAT_BATCHES, TARGETS, log_cases, stat_cases,  # noqa: E402
                           tally_cases, trace_cases, transport_cases)


def run(args, stdin=None):
    return subprocess.run([HARNESS] + [str(a) for a in args], input=stdin, capture_output=True,
                          text=True, check=True).stdout


def target_args(rho, els):
    a = [rho, len(els)]
    for Z, m, t in els:
        a += [Z, m, t]
    return a


def stopping_grid():
    e_in = "\n".join("%.17g" % e for e in ENERGIES) + "\n"
    out = np.zeros((len(IONS), len(TARGETS), len(ENERGIES)))
    for i, (Z1, m1) in enumerate(IONS):
        for j, (_, rho, els) in enumerate(TARGETS):
            txt = run(["stop", Z1, "%.17g" % m1] + target_args(rho, els), e_in)
            out[i, j] = [float(x) for x in txt.split()]
    np.savez_compressed(os.path.join(HERE, "stopping_grid.npz"), se=out, energies=ENERGIES,
                        ions=np.array(IONS), targets=np.array([t[0] for t in TARGETS]))


def average_table():
    rows = []
    for Z1, m1 in IONS:
        if m1 == 0.0:
            continue
        for _, rho, els in TARGETS:
            txt = run(["avg", Z1, "%.17g" % m1, 0.2] + target_args(rho, els))
            v = [float(x) for x in txt.split()]
            rows.append(v + [0.0] * (17 - len(v)))
    np.savez_compressed(os.path.join(HERE, "average.npz"), avg=np.array(rows))


def spec_file(prob):
    f = tempfile.NamedTemporaryFile("w", suffix=".spec", delete=False)
    f.write(prob.spec_text())
    f.close()
    return f.name


def traces(max_rows=1500, only=None):
    for name, prob in trace_cases().items():
        if only and name not in only:
            continue
        path = spec_file(prob)
        try:
            for h in (0, 1):
                txt = run(["trace", path, h, max_rows])
                rows = np.array([[float(x) for x in l.split()] for l in txt.splitlines()])
                np.savez_compressed(os.path.join(HERE, "trace_%s_h%d.npz" % (name, h)), rows=rows)
        finally:
            os.unlink(path)


def parse_log(txt, kinds):
    """Lines of one log mode -> rows [hist, kind, numbers...]; `kinds` maps a
    leading letter to a kind code (None: no letter)."""
    rows, h = [], -1
    for line in txt.splitlines():
        f = line.split()
        if not f:
            continue
        if f[0] == "H":
            h = int(f[1])
            continue
        if kinds:
            rows.append([h, kinds[f[0]]] + [float(x) for x in f[1:]])
        else:
            rows.append([h, 0] + [float(x) for x in f])
    return np.array(rows, np.float64).reshape(-1, 2 + (len(rows[0]) - 2 if rows else 0))


def logs():
    for name, (prob, vz) in log_cases().items():
        path = spec_file(prob)
        try:
            z = [str(v) for v in vz]
            d = parse_log(run(["log", path, "defects"]), {"V": 1, "I": 2, "S": 3, "R": 4})
            # TrimPhononOut prints recoils before their id is set (an
            # uninitialised int in the reference): the id column is dropped
            p = np.delete(parse_log(run(["log", path, "phonons"]), None), 9, axis=1)
            h = parse_log(run(["log", path, "history"]), None)
            m = run(["log", path, "vacmap"] + z).split("M", 1)[1].split()
        finally:
            os.unlink(path)
        np.savez_compressed(os.path.join(HERE, "log_%s.npz" % name), defects=d, phonons=p, history=h,
                            vacmap=np.array([int(v) for v in m], np.int32).reshape(20, 20, 3),
                            vacmap_Z=np.array(vz, np.int32))
        print(name, "defects", len(d), "phonons", len(p), "history", len(h), flush=True)


def parse_tally(txt):
    res = {}
    for line in txt.splitlines():
        k, *v = line.split()
        if len(v) == 1 and not k.startswith("h_"):
            res[k] = v[0]
        elif k in ("h_Eel", "h_Enuc", "h_E0", "h_end"):
            res[k] = [float(x) for x in v]
        else:
            res[k] = [int(x) for x in v]
    return res


def stat_fixtures(only=None):
    """Each case in STAT_BATCHES equal batches, run in parallel; the batches'
    depth histograms are kept apart (batch means give their covariance)."""
    for name, prob in stat_cases().items():
        if only and name not in only:
            continue
        assert prob.n_hist % STAT_BATCHES == 0
        m = prob.n_hist // STAT_BATCHES
        tmp = tempfile.mkdtemp()
        try:
            pkafile = None
            if not prob.primaries.uniform:
                pkafile = os.path.join(tmp, "pka.txt")
                with open(pkafile, "w") as f:
                    f.write(prob.primary_rows())
            specs = []
            for i in range(STAT_BATCHES):
                path = os.path.join(tmp, "b%d.spec" % i)
                with open(path, "w") as f:
                    f.write(prob.with_(hist0=i * m, n_hist=m).spec_text(pkafile) + "perhist 1\n")
                specs.append(path)
            with ThreadPoolExecutor(min(8, os.cpu_count() or 1)) as ex:
                parts = list(ex.map(lambda p: parse_tally(run(["tally", p])), specs))
        finally:
            for f in os.listdir(tmp):
                os.unlink(os.path.join(tmp, f))
            os.rmdir(tmp)
        out = {"n_hist": prob.n_hist, "seed": prob.seed, "batches": STAT_BATCHES}
        for k in ("h_vac", "h_coll"):
            out[k] = np.concatenate([np.array(p[k], np.int64) for p in parts])
        for k in ("h_Eel", "h_Enuc", "h_E0"):
            out[k] = np.concatenate([np.array(p[k], np.float64) for p in parts])
        for k in ("range_hist", "lat_y_hist", "lat_z_hist"):
            out[k] = np.sum([np.array(p[k], np.int64) for p in parts], axis=0)
        for k in ("vac_hist", "repl_hist"):
            out[k + "_batches"] = np.array([p[k] for p in parts], np.int32)
        for k in ("n_coll", "n_flight", "vac", "repl", "subst", "inter", "lost", "vacuum", "disp",
                  "range_n"):
            out[k] = np.int64(sum(int(p[k]) for p in parts))
        for k in ("E0", "Eel", "Enuc", "Eunf", "Elost", "Edrop", "Evac", "range_sum", "range_sum2"):
            out[k] = np.float64(sum(float(p[k]) for p in parts))
        np.savez_compressed(os.path.join(HERE, "stat_%s.npz" % name), problem=prob.name, **out)
        print(name, "n_hist", prob.n_hist, "coll/hist", out["n_coll"] / prob.n_hist,
              "vac/ion", out["vac"] / prob.n_hist, flush=True)


def tallies(quick, only=None):
    for name, (prob, nproc) in tally_cases(quick).items():
        if only and name not in only:
            continue
        chunk = (prob.n_hist + nproc - 1) // nproc
        parts = []
        specs = []
        for i in range(nproc):
            h0 = i * chunk
            n = min(chunk, prob.n_hist - h0)
            if n <= 0:
                continue
            specs.append(spec_file(prob.with_(hist0=h0, n_hist=n)))
        try:
            with ThreadPoolExecutor(len(specs)) as ex:
                parts = list(ex.map(lambda p: parse_tally(run(["tally", p])), specs))
        finally:
            for p in specs:
                os.unlink(p)
        tot = {}
        for k in parts[0]:
            if isinstance(parts[0][k], list):
                if k == "h_vac":
                    tot[k] = sum((p[k] for p in parts), [])
                else:
                    tot[k] = [int(x) for x in np.sum([p[k] for p in parts], axis=0)]
            elif "." in parts[0][k] or "e" in parts[0][k]:
                tot[k] = float(sum(float(p[k]) for p in parts))
            else:
                tot[k] = int(sum(int(p[k]) for p in parts))
        tot["n_hist"] = prob.n_hist
        tot["seed"] = prob.seed
        tot["problem"] = prob.name
        with open(os.path.join(HERE, "tally_%s.json" % name), "w") as f:
            json.dump(tot, f)
        print(name, "n_hist", prob.n_hist, "coll", tot["n_coll"], "vac/ion",
              tot["vac"] / prob.n_hist, flush=True)


RUNMYTRIM_REF = os.path.join(ROOT, "oracle", "_ref", "runmytrim_ref")
REF = "/root/reference"
# case -> (deck under the reference tree, overrides of its options / ion blocks)
RUNMYTRIM_CASES = {
    "cu_on_cu_150keV": ("tests/json/cu_on_cu.json", {"number": 5}),
    "cu_on_cu_150keV_threads": ("tests/json/cu_on_cu.json", {"number": 6, "threads": 3}),
    "cu_on_cu_150keV_evac": ("tests/json/cu_on_cu.json", {"number": 3, "type": "vacenergycount"}),
    "cu_on_cu_150keV_range": ("validation/cu_on_cu/cu_on_cu.json", {"number": 200, "threads": 1}),
    "xe_on_uo2_scale10": ("tests/json/xe_on_uo2_scale10.json", {"number": 1}),
    "cu_on_cu_vac_1keV": ("validation/vacancy_count/cu_on_cu_vac.json", {}),
    "si_on_c_vac": ("validation/vacancy_count/si_on_c_vac.json", {"number": 200}),
    "cu_on_cu_range_1keV": ("validation/vacancy_count/cu_on_cu_range.json", {"number": 2000}),
}


def strip_comments(text):
    out = []
    for line in text.splitlines():
        i = line.find("//")
        out.append(line if i < 0 else line[:i])
    return "\n".join(out)


def runmytrim_spec(deck, base):
    m = deck["mytrim"]
    opt, ion, out = m.get("options", {}), m["ion"], m["output"]
    lines = ["type %s" % out["type"], "base %s" % base, "seed %d" % opt["seed"],
             "threads %d" % opt.get("threads", 1),
             # the reference leaves primaries_only uninitialised when the deck omits it
             "primaries_only %d" % int(bool(out.get("primaries_only", False)))]
    if "scale" in opt:
        lines.append("scale %r" % float(opt["scale"]))
    ion_line = "ion %d %r %r %d" % (ion["Z"], float(ion["mass"]), float(ion["energy"]), ion["number"])
    if "final_energy" in ion:
        ion_line += " %r" % float(ion["final_energy"])
    lines.append(ion_line)
    for L in m["sample"]["layers"]:
        lines.append("layer %r %r" % (float(L["thickness"]), float(L["rho"])))
        for e in L["elements"]:
            lines.append("el %d %r %r %s %s" % (e["Z"], float(e["mass"]), float(e["fraction"]),
                                                repr(float(e["edisp"])) if "edisp" in e else "-",
                                                repr(float(e["elbind"])) if "elbind" in e else "-"))
    return "\n".join(lines) + "\n"


def runmytrim_fixtures():
    for name, (rel, over) in RUNMYTRIM_CASES.items():
        with open(os.path.join(REF, rel)) as f:
            deck = json.loads(strip_comments(f.read()))
        m = deck["mytrim"]
        for k, v in over.items():
            if k in ("number",):
                m["ion"][k] = v
            elif k in ("threads", "seed", "scale"):
                m.setdefault("options", {})[k] = v
            elif k == "type":
                m["output"]["type"] = v
        m["output"]["base"] = name
        dst = os.path.join(HERE, "runmytrim", name)
        os.makedirs(dst, exist_ok=True)
        with tempfile.TemporaryDirectory() as tmp:
            spec = os.path.join(tmp, "spec.txt")
            with open(spec, "w") as f:
                f.write(runmytrim_spec(deck, name))
            res = subprocess.run([RUNMYTRIM_REF, spec], cwd=tmp, capture_output=True, text=True, check=True)
            totals = dict((k, float(v)) for k, v in (l.split() for l in res.stdout.splitlines()))
            suffix = {"vaccount": "_vac.dat", "vacenergycount": "_evac.dat", "range": "_ranges.dat"}
            outname = name + suffix[m["output"]["type"]]
            with open(os.path.join(tmp, outname)) as f:
                data = f.read()
        with open(os.path.join(dst, outname), "w") as f:
            f.write(data)
        with open(os.path.join(dst, "deck.json"), "w") as f:
            json.dump(deck, f, indent=1)
        totals["source"] = rel
        totals["npka"] = m["ion"]["number"]
        with open(os.path.join(dst, "totals.json"), "w") as f:
            json.dump(totals, f, indent=1)
        print(name, totals, flush=True)


RUNSTOPPING_REF = os.path.join(ROOT, "oracle", "_ref", "runstopping_ref")
# case -> deck under the reference tree (its "stopping" block)
RUNSTOPPING_CASES = {
    "h_on_fe": "validation/h_on_fe/input.json",
    "he_on_fe": "validation/he_on_fe/input.json",
    "c_on_w": "validation/c_on_w/input.json",
}
# extra decks written here: the array and single-number energy forms
RUNSTOPPING_EXTRA = {
    "xe_on_uo2_list": {"stopping": {"ion": {"Z": 54, "mass": 131.904,
                                            "energy": [10, 1000, 30000, 1e6, 1e8]},
                                    "material": {"rho": 10.97, "elements": [
                                        {"Z": 92, "mass": 238.03, "fraction": 1},
                                        {"Z": 8, "mass": 15.999, "fraction": 2}]}}},
    "u_on_u_single": {"stopping": {"ion": {"Z": 92, "mass": 238.03, "energy": 2e6},
                                   "material": {"rho": 19.05, "elements": [
                                       {"Z": 92, "mass": 238.03, "fraction": 1}]}}},
}


def runstopping_fixtures():
    decks = {}
    for name, rel in RUNSTOPPING_CASES.items():
        with open(os.path.join(REF, rel)) as f:
            decks[name] = json.loads(strip_comments(f.read()))
    decks.update(RUNSTOPPING_EXTRA)
    for name, deck in decks.items():
        st = deck["stopping"]
        lines = ["material %r" % float(st["material"]["rho"])]
        for e in st["material"]["elements"]:
            lines.append("el %d %r %r" % (e["Z"], float(e["mass"]), float(e["fraction"])))
        lines.append("ion %d %r" % (st["ion"]["Z"], float(st["ion"]["mass"])))
        E = st["ion"]["energy"]
        if isinstance(E, dict):
            mode = "step" if "step" in E else "mult"
            lines.append("%s %r %r %r" % (mode, float(E["begin"]), float(E["end"]), float(E[mode])))
        else:
            lines.append("energies " + " ".join(repr(float(x)) for x in (E if isinstance(E, list) else [E])))
        dst = os.path.join(HERE, "runstopping", name)
        os.makedirs(dst, exist_ok=True)
        with tempfile.TemporaryDirectory() as tmp:
            spec = os.path.join(tmp, "spec.txt")
            with open(spec, "w") as f:
                f.write("\n".join(lines) + "\n")
            out = subprocess.run([RUNSTOPPING_REF, spec], capture_output=True, text=True, check=True).stdout
        with open(os.path.join(dst, "deck.json"), "w") as f:
            json.dump({"stopping": st}, f, indent=1)
        with open(os.path.join(dst, "stdout.txt"), "w") as f:
            f.write(out)
        print(name, len(out.splitlines()), "lines", flush=True)


def transport():
    """Layer transport (ref_harness log ... transport: the bookkeeping of
    apps/mytrim_dynamic.C:160-178 with fixed layers) of transport_cases()."""
    res = {}
    for name, prob in transport_cases().items():
        path = spec_file(prob)
        try:
            out = run(["log", path, "transport"])
        finally:
            os.unlink(path)
        res[name] = [[int(v) for v in line.split()[1:]] for line in out.splitlines() if line.startswith("T ")]
        print(name, "transport", res[name], flush=True)
    with open(os.path.join(HERE, "transport.json"), "w") as f:
        json.dump(res, f, indent=0)


def dynamic():
    """SampleDynamic: the layer updates of addAtomsToLayer (ref_harness
    dynamic) and the drain loop of tests/native/dynamic_trim.cpp built against
    the reference (oracle/_ref/dynamic_trim), 6 Xe primaries."""
    out = subprocess.run([HARNESS, "dynamic"], capture_output=True, text=True, check=True).stdout
    with open(os.path.join(HERE, "dynamic_layers.txt"), "w") as f:
        f.write(out)
    drv = os.path.join(os.path.dirname(HARNESS), "dynamic_trim")
    out = subprocess.run([drv, "6"], capture_output=True, text=True, check=True).stdout
    with open(os.path.join(HERE, "dynamic_trim.txt"), "w") as f:
        f.write(out)
    print("dynamic", len(out.splitlines()), "lines", flush=True)


def main():
    if not os.path.exists(HARNESS):
        sys.exit("build the reference harness first: make -C oracle ref")
    quick = "--quick" in sys.argv
    if "--runmytrim" in sys.argv:
        runmytrim_fixtures()
        runstopping_fixtures()
        return
    only = [a.split("=", 1)[1] for a in sys.argv if a.startswith("--tally=")]
    only_tr = [a.split("=", 1)[1] for a in sys.argv if a.startswith("--trace=")]
    if "--logs" in sys.argv:
        logs()
        return
    stat = [a for a in sys.argv if a.startswith("--stat")]
    if stat:
        stat_fixtures([a.split("=", 1)[1] for a in stat if "=" in a] or None)
        return
    if "--dynamic" in sys.argv:
        dynamic()
        transport()
        return
    if only or only_tr:  # regenerate the named tally / trace fixtures only
        if only_tr:
            traces(only=only_tr)
        if only:
            tallies(quick, only)
        return
    stopping_grid()
    average_table()
    traces()
    logs()
    dynamic()
    transport()
    tallies(quick)
    stat_fixtures()


if __name__ == "__main__":
    main()
