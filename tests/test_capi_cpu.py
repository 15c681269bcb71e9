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
    small = lib.bta_ws_bytes(2, 64, 4, 0)
    big = lib.bta_ws_bytes(200, 2000, 6, 0)
    assert big > small > 2 * 64 * 64 * 8
    assert big >= 200 * 32 * 64 * 64 * 8


def test_workspace_bytes_count_every_reservation():
    """bta_ws_bytes counts what a workspace holds: the block inverses W_i
    (n b^2 doubles) with a solve reservation, plus G_i / Ga_i (n b^2 +
    n max(a,1) b) with a selected-inversion reservation, and the persistent
    panel's exchange buffers ((2 nT + nTa) nT + 2 nT tiles of 64x64 doubles
    plus one flag word per tile and 3 nT more) for b > 64."""
    lib = _lib.load()
    n, b, a = 200, 2000, 6
    nT = (b + 63) // 64
    base = lib.bta_ws_bytes(n, b, a, 0)
    solve = lib.bta_ws_bytes(n, b, a, _lib.WS_SOLVE)
    sel = lib.bta_ws_bytes(n, b, a, _lib.WS_SOLVE | _lib.WS_SELINV)
    assert solve - base == 8 * n * b * b
    assert sel - solve == 8 * (n * b * b + n * a * b)
    # a = 0: G_a still has one row per block
    assert lib.bta_ws_bytes(n, b, 0, 6) - lib.bta_ws_bytes(n, b, 0, 2) == 8 * (n * b * b + n * b)
    flow = ((2 * nT + 1) * nT + 2 * nT) * 4096 * 8 + ((2 * nT + 1) * nT + 3 * nT) * 8
    tiles = n * nT * 4096 * 8
    assert base >= tiles + flow
    assert base - (tiles + flow) < 3 * b * b * 8 + 10 * 4096 * 8 + 64 * n * nT  # scratch and flags only
    # b <= 64 runs the per-column panel: no exchange buffers
    flow1 = 5 * 4096 * 8  # exchange buffers of a one-column panel
    assert lib.bta_ws_bytes(10, 64, 4, 0) < 10 * 4096 * 8 + 3 * 64 * 64 * 8 + 4096 * 8 + flow1 / 2


def test_scatter_key_uses_every_index():
    """Two sparsity patterns with the same nnz and the same first and last
    entries get different scatter-map keys; the same pattern the same key."""
    import numpy as np
    from pyinla_amd.solver import SerinvSolver
    rows = np.array([0, 1, 2, 3, 4, 5, 6, 7])
    cols = np.array([0, 0, 1, 2, 3, 4, 5, 7])
    rows2 = rows.copy()
    cols2 = cols.copy()
    cols2[4] = 2  # differs in the middle only
    k1 = SerinvSolver._pattern_key(rows, cols)
    assert k1 == SerinvSolver._pattern_key(rows.copy(), cols.copy())
    assert k1 != SerinvSolver._pattern_key(rows2, cols2)
    assert k1 != SerinvSolver._pattern_key(cols, rows)
