"""The C ABI: the library builds for gfx950, loads, and exports every declared entry.

No compute call needs a GPU here; these run on the CPU.
"""

import re
from pathlib import Path

from solprop_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "solprop_mi355x.h"


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sp_[a-z0-9_]+)\s*\(", text)))


def test_library_is_built_for_gfx950():
    assert _lib.LIB_PATH.exists(), "run python -m solprop_amd.build"
    data = _lib.LIB_PATH.read_bytes()
    assert b"gfx950" in data


def test_every_declared_symbol_is_exported():
    lib = _lib.lib()
    declared = declared_symbols()
    assert declared, "no entry points parsed from the header"
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_lib.exported_symbols())


def test_version_string():
    assert b"gfx950" in _lib.lib().sp_version()


def test_refusals_are_values_not_crashes():
    import ctypes as C

    lib = _lib.lib()
    handle = C.c_void_p()
    msg = C.create_string_buffer(256)
    assert lib.sp_workspace_create(None, C.byref(handle), msg, 256) == _lib.SP_VALUE
    assert b"problem" in msg.value
