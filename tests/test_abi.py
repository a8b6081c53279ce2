"""The C-ABI library (CPU only: loads it, no GPU calls)."""
import ctypes as C
import os
import re

import pytest

import timbir_amd as tb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in ("mbir4d.h", "timbir_dev.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b([a-z_][a-z0-9_]*)\s*\(", text):
            name = m.group(1)
            if name.startswith("tb_") or name == "reconstruct":
                names.append(name)
    return sorted(set(names))


def test_library_exports_every_declared_entry_point():
    lib = tb.lib()
    names = declared_functions()
    assert "reconstruct" in names and "tb_stage_create" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(set(tb.EXPORTS)) == names


def test_options_round_trip_without_gpu():
    o = tb.get_options()
    assert o.max_iterations == 1000
    tb.set_options(order_seed=7)
    assert tb.get_options().order_seed == 7
    tb.set_options(order_seed=0)


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(tb, "_LIB", None)
    monkeypatch.setattr(tb, "LIB_PATH", "/nonexistent/libtimbir_hip.so")
    with pytest.raises(tb.TimbirError):
        tb.lib()
