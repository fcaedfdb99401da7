"""The C-ABI library builds, loads on the CPU and exports every symbol that
include/phaser_mi355x.h declares (no compute calls: there is no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for name in os.listdir(os.path.join(ROOT, "include")):
        if name.endswith(".h"):
            text = open(os.path.join(ROOT, "include", name)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            syms |= set(re.findall(r"^\s*(?:int|int64_t|void)\s+(phx_\w+)\s*\(", text, flags=re.M))
    return syms


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert {"phx_get_cutouts_c64", "phx_add_cutouts_f32", "phx_fft2_c64", "phx_lsqml_backward",
            "phx_farfield", "phx_epie_backward"} <= syms


def test_library_exports_every_declared_symbol():
    from phaser_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built (run python -m phaser_amd.build)")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in sorted(declared_symbols()) if not hasattr(lib, s)]
    assert not missing, missing
    # and the Python binding covers them all
    assert declared_symbols() <= set(_lib.SIGNATURES)
    assert lib.phx_red_doubles(8, 64, 4) > 0


def test_missing_library_fails_loudly(tmp_path):
    from phaser_amd import _lib
    with pytest.raises(RuntimeError, match="not built"):
        _lib.load.__wrapped__(str(tmp_path / "nope.so")) if hasattr(_lib.load, "__wrapped__") else \
            _lib_load_fresh(str(tmp_path / "nope.so"))


def _lib_load_fresh(path):
    from phaser_amd import _lib
    saved = _lib._lib
    _lib._lib = None
    try:
        return _lib.load(path)
    finally:
        _lib._lib = saved
