"""CPU-side checks of the C ABI: the library loads and exports every declared symbol, and
the header, the ctypes binding and the library agree.  No compute call is made here."""
import ctypes
import os
import re

import pytest

from bd_box_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bdbox_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(bdh_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_binding_exports():
    assert declared_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libbdbox_hip.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_loads_without_gpu():
    L = _lib.load()
    # pure host helpers only
    assert L.bdh_seed_state(12345) == ((12345 << 16) | 0x330E)
    assert L.bdh_last_error() is not None


def test_config_layout_matches_header():
    # field order of bdh_config in the header == ctypes structure order
    text = open(HEADER).read()
    body = text[text.index("typedef struct bdh_config"):text.index("} bdh_config;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S).split("{", 1)[1]
    names = re.findall(r"\b([A-Za-z_][A-Za-z_0-9]*)\s*[,;]", body)
    fields = [f for f, _ in _lib.Config._fields_]
    assert names == fields
