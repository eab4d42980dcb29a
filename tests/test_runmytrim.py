"""runmytrim, the JSON front end (apps/runmytrim.C): input handling on the CPU;
on the GPU, its output files against the ones the reference's own driver loop
and output modules write for the same decks (tests/golden/runmytrim), and the
batched mode against them statistically."""
from __future__ import annotations

import json
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mytrim_amd", "host", "_build", "runmytrim")
GOLD = os.path.join(ROOT, "tests", "golden", "runmytrim")
CASES = sorted(os.listdir(GOLD))


def run(deck_text, cwd, *args, timeout=600):
    if not os.path.exists(EXE):
        pytest.fail("runmytrim not built: run __graft_entry__.build()")
    return subprocess.run([EXE, *args], input=deck_text, cwd=cwd, capture_output=True, text=True,
                          timeout=timeout)


DECK = """{ "mytrim" : {
  "options": { "seed": 2344 },   // comment after a value
  /* block
     comment */
  "ion" : { "Z": 29, "mass": 63.546, "energy": 1000, "number": 2 },
  "sample": { "layers": [ { "thickness": 100, "rho": 8.92, "elements": [
        { "Z": 29, "mass": 63.546, "fraction": 1 } ] } ] },
  "output" : { "base": "x", "type": "%s" }
} }"""


# ------------------------------------------------------------------ CPU: input handling
@pytest.mark.parametrize("text,msg", [
    ("{", "JSON parse error"),
    ('{"other": {}}', "No 'mytrim' top level block"),
    ('{"mytrim": {"options": {"seed": 1}}}', "Must specify a 'sample' block"),
    ('{"mytrim": {"options": {"seed": 1}, "sample": {"layers": 3}}}', "sample.layers must be an array"),
    ('{"mytrim": {"options": {"seed": 1}, "sample": {"layers": [{"rho": 1}]}}}', "No 'thickness' found for layer 0"),
    ('{"mytrim": {"options": {"seed": 1}, "sample": {"layers": []}}}', "Must specify an 'output' block"),
    (DECK % "bogus", "Unknown output type bogus"),
    (DECK.replace('"Z": 29, "mass": 63.546, "energy"', '"mass": 63.546, "energy"') % "vaccount",
     "Missing 'Z' in ion block"),
    (DECK.replace('"fraction": 1', '"frac": 1') % "vaccount", "Missing 'fraction' in element 0 in layer 0"),
])
def test_input_errors(tmp_path, text, msg):
    r = run(text, str(tmp_path))
    assert r.returncode == 1
    assert msg in r.stderr, r.stderr


def test_arguments_rejected(tmp_path):
    r = run(DECK % "vaccount", str(tmp_path), "input.json")
    assert r.returncode == 1 and "via stdin" in r.stderr
    r = run(DECK % "range", str(tmp_path), "--batch")
    assert r.returncode == 1 and "--batch supports output type vaccount only" in r.stderr


def test_fixture_decks_parse_like_python_json():
    """Every committed deck is valid JSON; the reference decks with comments
    are exercised by the GPU tests below."""
    for c in CASES:
        with open(os.path.join(GOLD, c, "deck.json")) as f:
            d = json.load(f)
        assert d["mytrim"]["output"]["base"] == c


# ------------------------------------------------------------------ GPU: against the reference
def out_name(case, deck):
    t = deck["mytrim"]["output"]["type"]
    return case + {"vaccount": "_vac.dat", "vacenergycount": "_evac.dat", "range": "_ranges.dat"}[t]


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_output_files_match_reference(tmp_path, case):
    with open(os.path.join(GOLD, case, "deck.json")) as f:
        text = f.read()
    deck = json.loads(text)
    with open(os.path.join(GOLD, case, "totals.json")) as f:
        tot = json.load(f)
    print("case", case, flush=True)
    r = run(text, str(tmp_path))
    assert r.returncode == 0, r.stderr
    name = out_name(case, deck)
    with open(os.path.join(GOLD, case, name)) as f:
        want = f.read()
    with open(os.path.join(str(tmp_path), name)) as f:
        got = f.read()
    assert got == want
    vpi = float(re.search(r"Vacancies/ion: (\S+)", r.stderr).group(1))
    assert vpi == pytest.approx(tot["vacancies"] / tot["npka"], rel=1e-5)
    eel = float(re.search(r"Electronic energy loss/ion: (\S+)", r.stderr).group(1))
    assert eel == pytest.approx(tot["EelTotal"] / tot["npka"], rel=1e-5)


@pytest.mark.gpu
def test_batch_mode_statistics(tmp_path):
    """--batch (one launch, Philox streams): vacancies/ion and the mean
    vacancy depth of 1 keV Cu -> Cu against the reference loop's 1000 ions."""
    case = "cu_on_cu_vac_1keV"
    with open(os.path.join(GOLD, case, "deck.json")) as f:
        deck = json.load(f)
    with open(os.path.join(GOLD, case, "totals.json")) as f:
        tot = json.load(f)
    ref = np.loadtxt(os.path.join(GOLD, case, case + "_vac.dat"))
    deck["mytrim"]["ion"]["number"] = 50000
    r = run(json.dumps(deck), str(tmp_path), "--batch")
    assert r.returncode == 0, r.stderr
    print(r.stderr)
    got = np.loadtxt(os.path.join(str(tmp_path), case + "_vac.dat"))
    vpi = got[:, 1].sum() / 50000
    vpi_ref = tot["vacancies"] / tot["npka"]
    assert vpi == pytest.approx(vpi_ref, rel=0.05)
    depth = (got[:, 0] * got[:, 1]).sum() / got[:, 1].sum()
    depth_ref = (ref[:, 0] * ref[:, 1]).sum() / ref[:, 1].sum()
    assert depth == pytest.approx(depth_ref, rel=0.05, abs=0.5)
