"""Tile-shape autotuner for the FP64-MFMA GEMMs of the BTA solver.

For each GEMM class (block updates of the factorization; block inverses +
selected inversion) the candidate output-tile edges (64, 128) are timed on the
solver's own matrix with HIP events, the blocks being restored from a device
copy before every trial, and the fastest is kept in the solver's workspace.
Results are cached per (n, b, a) in memory and, if a path is given, in a JSON
file.
"""
from __future__ import annotations

import json
import os

import torch

from . import _lib

OPT_TILE_UPDATE = 1
OPT_TILE_GROUPED = 2

_cache: dict = {}


def _key(s, sparsity):
    a = s.arrowhead_blocksize if sparsity == "bta" else 0
    return f"{s.n_diag_blocks}x{s.diagonal_blocksize}x{a}"


def _blocks(s):
    return [t for t in (s.A_diagonal_blocks, s.A_lower_diagonal_blocks, s.A_arrow_bottom_blocks,
                        s.A_arrow_tip_block) if t is not None]


def _time(fn, reps):
    best = float("inf")
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    return best


def apply(solver, choice: dict) -> None:
    lib = _lib.load()
    _lib.check(lib.bta_ws_set_option(solver._ws, OPT_TILE_UPDATE, int(choice.get("update", 0))), "set_option")
    _lib.check(lib.bta_ws_set_option(solver._ws, OPT_TILE_GROUPED, int(choice.get("grouped", 0))), "set_option")


def autotune(solver, sparsity: str = "bta", candidates=(64, 128), reps: int = 2, cache_path: str | None = None,
             selected_inversion: bool = True) -> dict:
    """Pick the tile edge per GEMM class for ``solver``'s shape.

    The solver's buffers must hold the (unfactored) matrix; they hold it again
    on return.  Returns {"update": edge, "grouped": edge, "times_ms": {...}}.
    """
    key = _key(solver, sparsity)
    if cache_path and os.path.exists(cache_path) and not _cache:
        with open(cache_path) as f:
            _cache.update(json.load(f))
    if key in _cache:
        apply(solver, _cache[key])
        return _cache[key]
    saved = [t.clone() for t in _blocks(solver)]

    def restore():
        for dst, src in zip(_blocks(solver), saved):
            dst.copy_(src)

    times = {}
    best_u, best_g = None, None
    for tu in candidates:
        apply(solver, {"update": tu, "grouped": 0})

        def run_f():
            solver.factorize(sparsity)
        restore()
        run_f()  # warm (builds plans, captures graphs)
        t = _time(lambda: (restore(), run_f()), reps)
        times[f"update_{tu}"] = t
        if best_u is None or t < times[f"update_{best_u}"]:
            best_u = tu
    if selected_inversion:
        for tg in candidates:
            apply(solver, {"update": best_u, "grouped": tg})
            restore()
            solver.factorize(sparsity)
            solver.selected_inversion(sparsity)  # warm
            torch.cuda.synchronize()

            def run_si():
                restore()
                solver.factorize(sparsity)
                solver.selected_inversion(sparsity)
            t = _time(run_si, reps)
            times[f"grouped_{tg}"] = t
            if best_g is None or t < times[f"grouped_{best_g}"]:
                best_g = tg
    choice = {"update": best_u, "grouped": best_g or 0, "times_ms": times}
    apply(solver, choice)
    restore()
    _cache[key] = choice
    if cache_path:
        with open(cache_path, "w") as f:
            json.dump(_cache, f, indent=1)
    del saved
    return choice
