"""CPU checks of the C ABI: the library loads and exports every declared symbol.

No compute entry point is called here (no GPU); only the pure host functions
bta_version and bta_ws_bytes are.
"""
import ctypes as C
import re
from pathlib import Path

import pytest

from pyinla_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def _declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = h.read_text()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names.update(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(bta_\w+)\s*\(", txt, flags=re.M))
    return sorted(names)


def test_header_declares_entry_points():
    syms = _declared_symbols()
    assert "bta_potrf" in syms and "bta_potrs" in syms and "bta_potri" in syms
    assert len(syms) >= 8


def test_library_exports_every_declared_symbol():
    if not _lib.LIB_PATH.exists():
        pytest.fail("libpyinla_bta.so is not built (run python -m pyinla_amd.build_lib)")
    lib = C.CDLL(str(_lib.LIB_PATH))
    missing = [s for s in _declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(_declared_symbols()) <= set(_lib.SIGNATURES)


def test_host_only_functions():
    lib = _lib.load()
    assert lib.bta_version() >= 100
    # workspace bytes grow with the problem and include the tile inverses
    small = lib.bta_ws_bytes(2, 64, 4)
    big = lib.bta_ws_bytes(200, 2000, 6)
    assert big > small > 2 * 64 * 64 * 8
    assert big >= 200 * 32 * 64 * 64 * 8
