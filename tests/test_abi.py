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
