"""CPU checks of the C ABI library: it loads and exports every symbol that
include/mytrim_gpu.h declares (no compute call; there is no GPU here)."""
import ctypes
import os
import re
import subprocess

import pytest

from mytrim_amd import gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "mytrim_gpu.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(mt_[a-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(gpu.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(gpu.LIB_PATH):
        pytest.skip("library not built")
    L = ctypes.CDLL(gpu.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", gpu.LIB_PATH], capture_output=True,
                         text=True).stdout
    for name in declared_symbols():
        assert re.search(r"\bT %s$" % name, out, re.M), name


def test_version_string():
    if not os.path.exists(gpu.LIB_PATH):
        pytest.skip("library not built")
    assert b"gfx950" in gpu.lib().mt_version()


# ctypes mirror -> C struct of include/mytrim_gpu.h
STRUCTS = {"Target": "mt_target", "Primaries": "mt_primaries", "RunOpts": "mt_run_opts",
           "LogSpec": "mt_log", "Tallies": "mt_tallies", "Tuning": "mt_tuning",
           "SurrogateOpts": "mt_surrogate_opts", "Ion": "mt_ion", "Event": "mt_event"}


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes mirror has the size and field offsets the C compiler gives
    the header's struct (a field added on one side only shifts the rest)."""
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "mytrim_gpu.h"', "int main(void) {"]
    expect = []
    for py, c in STRUCTS.items():
        cls = getattr(gpu, py)
        lines.append('  printf("%%zu\\n", sizeof(%s));' % c)
        expect.append(ctypes.sizeof(cls))
        for f in cls._fields_:
            lines.append('  printf("%%zu\\n", offsetof(%s, %s));' % (c, f[0]))
            expect.append(getattr(cls, f[0]).offset)
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)],
                   check=True, capture_output=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got == expect
