"""ctypes binding of libpyinla_bta.so (signatures in include/pyinla_bta.h).

The library is the only compute path: if it is missing or fails to load, every
entry point raises immediately.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libpyinla_bta.so"

_P = C.c_void_p
_I = C.c_int
_D = C.c_double

# name -> (restype, argtypes); the authoritative list of exported symbols.
SIGNATURES = {
    "bta_version": (_I, []),
    "bta_ws_bytes": (C.c_size_t, [_I, _I, _I, _I]),
    "bta_ws_create": (_I, [_I, _I, _I, _I, _I, C.POINTER(_P)]),
    "bta_ws_destroy": (_I, [_P]),
    "bta_ws_reserve": (_I, [_P, _I]),
    "bta_ws_allocated": (C.c_size_t, [_P]),
    "bta_ws_prepare": (_I, [_P, _P, _P, _P, _P, _I, _I, _I, _I]),
    "bta_logdet": (_I, [_P, _P, _P, _I, _I, _I, _P, _P]),
    "bta_potrf": (_I, [_P, _P, _P, _P, _P, _I, _I, _I, _P]),
    "bta_potrf_noalloc": (_I, [_P, _P, _P, _P, _I, _I, _I, _P, _P, _P, _P]),
    "bta_ws_status": (_I, [_P, C.POINTER(C.c_int64), C.POINTER(_D), C.POINTER(_I), _P]),
    "bta_ws_logdet_ptr": (_P, [_P]),
    "bta_selftest_mfma": (_I, [_P, _P]),
    "bta_batch_create": (_I, [_I, _P, _P, _P, _P, _P, _P, _I, C.POINTER(_P)]),
    "bta_batch_potrf": (_I, [_P, _P]),
    "bta_batch_destroy": (_I, [_P]),
    "bta_ws_set_option": (_I, [_P, _I, _I]),
    "bta_ws_set_timing": (_I, [_P, _I]),
    "bta_ws_reset_timing": (_I, [_P]),
    "bta_ws_get_timing": (_I, [_P, _I, C.POINTER(_D), C.POINTER(C.c_int64)]),
    "bta_potrs": (_I, [_P, _P, _P, _P, _P, _I, _I, _I, _P, _I, _I, _I, _P]),
    "bta_potri": (_I, [_P, _P, _P, _P, _P, _I, _I, _I, _P]),
    "bta_block_inverses": (_I, [_P, _P, _I, _I, _P]),
    "bta_assemble_kron": (_I, [_P, _P, _P, _P, _I, _I, _I, _I, _P, _P, _P, _I, _I, _P, _P, _I, _P, _D, _P]),
    "bta_scatter_coo": (_I, [_P, _P, _P, _P, _P, _P, _P, C.c_int64, _D, _I, _P]),
    "bta_gather_coo": (_I, [_P, _P, _P, _P, _P, _P, C.c_int64, _P]),
    "bta_bilinear_coo": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P, C.c_int64, _I, _P, _I, _P, _P]),
    "bta_quadform_rows": (_I, [_P, _P, _P, _P, _I, _I, _I, _P, _P, _P, C.c_int64, _P, _P]),
    "bta_kron_contract": (_I, [_P, _P, _P, _I, _I, _I, _P, _P, _P, _I, _I, _I, _P, _P, _P]),
    "bta_gemm": (_I, [_P, _I, _I, _P, _I, _I, _P, _I, _I, _I, _I, _D, _D, _I, _P]),
    "bta_flow_schedule": (_I, [_I, _I, _I, _I, _I, _I, _P, C.c_int64, _P, C.POINTER(C.c_int64)]),
}

BTA_E = {-1: "invalid argument", -2: "shape does not match the workspace", -3: "in-kernel wait timed out",
         -4: "rocBLAS call failed", -5: "the workspace lacks the reservation this call needs",
         -6: "the shape exceeds the fixed device scratch of bta_potrf_noalloc"}

# workspace flags (include/pyinla_bta.h)
WS_GRAPH, WS_SOLVE, WS_SELINV = 1, 2, 4
RESERVE = {"factor": 0, "solve": WS_SOLVE, "selinv": WS_SOLVE | WS_SELINV}

_lib = None


class BTALibraryError(RuntimeError):
    pass


def load(path: Path | None = None):
    """Load the shared library once and bind every signature."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise BTALibraryError(
            f"{p} is missing: build it with `python -m pyinla_amd.build_lib` "
            "(the BTA solver has no CPU fallback)")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = BTA_E.get(rc, f"HIP error {rc}")
        raise BTALibraryError(f"{what} failed: {msg}")
