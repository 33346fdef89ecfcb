"""The C ABI library loads on a machine without a GPU and exports every entry point
include/rednative.h declares, with no compute call made."""
import os
import re

from conftest import ROOT


def declared_symbols():
    with open(os.path.join(ROOT, "include", "rednative.h")) as f:
        txt = f.read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(red_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "red_ctx_create" in syms and "red_step" in syms and "red_accel" in syms
    assert len(syms) >= 25


def test_library_exports_every_declared_symbol(native_mod):
    L = native_mod.load_library()
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_binding_covers_header(native_mod):
    missing = [s for s in declared_symbols() if s not in native_mod.SIGNATURES]
    assert not missing, missing


def test_version_and_null_context(native_mod):
    L = native_mod.load_library()
    assert L.red_version().startswith(b"rednative")
    assert L.red_last_error(None) == b"no context"
    assert L.red_step(None, None) == 1
