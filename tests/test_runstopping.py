"""runstopping (apps/runstopping.C): input handling on the CPU; on the GPU its
stdout against what the reference's loop prints for the same decks (the SRIM
validation decks of the reference and two decks for the array and
single-number energy forms), byte for byte."""
from __future__ import annotations

import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mytrim_amd", "host", "_build", "runstopping")
GOLD = os.path.join(ROOT, "tests", "golden", "runstopping")
CASES = sorted(os.listdir(GOLD))


def run(text, *args):
    if not os.path.exists(EXE):
        pytest.fail("runstopping not built: run __graft_entry__.build()")
    return subprocess.run([EXE, *args], input=text, capture_output=True, text=True, timeout=300)


BASE = {"stopping": {"ion": {"Z": 1, "mass": 1.008, "energy": {"begin": 1e4, "end": 1e5, "mult": 1.5}},
                     "material": {"rho": 7.87, "elements": [{"Z": 26, "mass": 55.845, "fraction": 1}]}}}


def deck(**edit):
    d = json.loads(json.dumps(BASE))
    for path, v in edit.items():
        node = d["stopping"]
        keys = path.split("__")
        for k in keys[:-1]:
            node = node[k]
        if v is None:
            del node[keys[-1]]
        else:
            node[keys[-1]] = v
    return json.dumps(d)


@pytest.mark.parametrize("text,msg", [
    ("[", "JSON parse error"),
    ('{"mytrim": {}}', "No 'stopping' top level block"),
    (deck(material=None), "Must specify a 'material' block"),
    (deck(material__rho=None), "Missing 'rho'"),
    (deck(ion__Z=None), "Missing 'Z' in ion block"),
    (deck(ion__energy="x"), "Missing or invalid 'energy' in ion block"),
    (deck(ion__energy={"end": 5.0, "mult": 2.0}), "Missing 'begin' in energy block"),
    (deck(ion__energy={"begin": 1.0, "end": 5.0}), "Specify either 'step' or 'mult' energy block"),
    (deck(ion__energy={"begin": 1.0, "end": 5.0, "mult": 1.0}), "'mult' must be larger than 1.0"),
])
def test_input_errors(text, msg):
    r = run(text)
    assert r.returncode == 1 and msg in r.stderr, r.stderr


def test_arguments_rejected():
    r = run(deck(), "input.json")
    assert r.returncode == 1 and "via stdin" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_stdout_matches_reference(case):
    with open(os.path.join(GOLD, case, "deck.json")) as f:
        text = f.read()
    with open(os.path.join(GOLD, case, "stdout.txt")) as f:
        want = f.read()
    print("case", case, flush=True)
    r = run(text)
    assert r.returncode == 0, r.stderr
    assert r.stdout == want
