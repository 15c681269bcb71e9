"""Autotuner of the BTA solver's launch options, timed on the solver's own matrix.

The search space is the workspace options of include/pyinla_bta.h that change
speed but not results beyond rounding:

  factorization       panel variant (persistent dataflow panel, or one launch
                      per 64-wide tile column), the near-pool divisor of the
                      persistent panel, whether it applies the block update
                      itself or leaves it to a grouped GEMM, the output tile
                      of the block-update GEMMs (64 / 128);
  selected inversion  output tile of the grouped GEMM (64 / 128), its 64x64
                      main loop (one or two LDS buffers), and whether the
                      Sigma_{i+1,i} products go to rocBLAS dgemm or stay in the
                      grouped kernel.

Coordinate descent: one option at a time, the others at their best so far;
each candidate is timed with HIP events (best of ``reps``, after one warm-up
run that builds its plans) on the solver's blocks, restored from a device copy
before every trial.  Results are cached per (n, b, a) in memory and, if a path
is given, in a JSON file, and applied to the workspace.  No reference
counterpart: the reference runs Serinv with fixed settings.
"""
from __future__ import annotations

import json
import os

import torch

from . import _lib

OPT_TILE_UPDATE = 1
OPT_TILE_GROUPED = 2
OPT_PANEL = 3
OPT_FLOW_NEAR = 8
OPT_BLAS = 7
OPT_GEMM_KERNEL = 9
OPT_FLOW_FOLD = 10

# (name, option, candidates); DEFAULTS holds the library defaults
FACTOR_SPACE = [("panel", OPT_PANEL, [2, 1]), ("near", OPT_FLOW_NEAR, [16, 8, 24]), ("fold", OPT_FLOW_FOLD, [1, 2]),
                ("update", OPT_TILE_UPDATE, [64, 128])]
SELINV_SPACE = [("grouped", OPT_TILE_GROUPED, [64, 128]), ("gemm_kernel", OPT_GEMM_KERNEL, [0, 1]),
                ("blas", OPT_BLAS, [1, 0])]
DEFAULTS = {"panel": 0, "near": 16, "fold": 0, "update": 0, "grouped": 0, "gemm_kernel": 0, "blas": 1}
OPTION = {name: opt for name, opt, _ in FACTOR_SPACE + SELINV_SPACE}

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
    """Set the workspace options of ``choice`` (missing names: library defaults)."""
    lib = _lib.load()
    for name, opt in OPTION.items():
        _lib.check(lib.bta_ws_set_option(solver._ws, opt, int(choice.get(name, DEFAULTS[name]))), "set_option")


def autotune(solver, sparsity: str = "bta", reps: int = 2, cache_path: str | None = None,
             selected_inversion: bool = True, factor_space=None, selinv_space=None) -> dict:
    """Pick the launch options for ``solver``'s shape.

    The solver's buffers must hold the (unfactored) matrix; they hold it again
    on return.  Returns {option name: value, ..., "times_ms": {candidate: ms}}.
    """
    key = _key(solver, sparsity)
    if cache_path and os.path.exists(cache_path) and not _cache:
        with open(cache_path) as f:
            _cache.update(json.load(f))
    if key in _cache:
        apply(solver, _cache[key])
        return _cache[key]
    fspace = FACTOR_SPACE if factor_space is None else factor_space
    sspace = SELINV_SPACE if selinv_space is None else selinv_space
    if solver.diagonal_blocksize <= 64:  # one tile column: no persistent panel to tune
        fspace = [d for d in fspace if d[0] not in ("panel", "near", "fold")]
    saved = [t.clone() for t in _blocks(solver)]

    def restore():
        for dst, src in zip(_blocks(solver), saved):
            dst.copy_(src)

    def run_f():
        restore()
        solver.factorize(sparsity)

    def run_si():
        restore()
        solver.factorize(sparsity)
        solver.selected_inversion(sparsity)

    choice = dict(DEFAULTS)
    times = {}

    def descend(space, run):
        for name, _opt, cands in space:
            best_v, best_t = None, float("inf")
            for v in cands:
                trial = dict(choice)
                trial[name] = v
                apply(solver, trial)
                run()  # warm: plans, schedules, graphs
                torch.cuda.synchronize()
                t = _time(run, reps)
                times[f"{name}={v}"] = t
                if t < best_t:
                    best_v, best_t = v, t
            choice[name] = best_v

    descend(fspace, run_f)
    if selected_inversion:
        descend(sspace, run_si)
    choice["times_ms"] = times
    apply(solver, choice)
    restore()
    _cache[key] = choice
    if cache_path:
        with open(cache_path, "w") as f:
            json.dump(_cache, f, indent=1)
    del saved
    return choice
