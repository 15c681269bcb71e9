"""Batched factorization of several solvers' buffers in one launch sequence.

All problems advance together: every block step / tile column is one launch
over every problem that has it, and the whole sequence replays as one
hipGraph.  Problems may differ in (n, b, a).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from .solver import SerinvSolver, _check_sparsity, _stream_ptr


class BatchedFactorization:
    def __init__(self, solvers: list[SerinvSolver], sparsities: list[str], use_graph: bool = True):
        if len(solvers) != len(sparsities) or not solvers:
            raise ValueError("need one sparsity per solver")
        self.lib = _lib.load()
        self.solvers = list(solvers)
        self.sparsities = list(sparsities)
        self.device = solvers[0].device
        q = len(solvers)
        P = C.c_void_p
        ws = (P * q)()
        D, L, B, T = (P * q)(), (P * q)(), (P * q)(), (P * q)()
        a = (C.c_int * q)()
        for i, (s, sp) in enumerate(zip(solvers, sparsities)):
            _check_sparsity(sp)
            if s.device != self.device:
                raise ValueError("all solvers must live on one device")
            Dt, Lt, Bt, Tt, ai = s._blocks(sp)
            ws[i] = s._ws
            D[i] = Dt.data_ptr()
            L[i] = Lt.data_ptr() if Lt.numel() else None
            B[i] = Bt.data_ptr() if Bt is not None else None
            T[i] = Tt.data_ptr() if Tt is not None else None
            a[i] = ai
        h = P()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.bta_batch_create(q, ws, D, L, B, T, a, int(use_graph), C.byref(h)),
                       "bta_batch_create")
        self._h = h

    def factorize(self) -> None:
        with torch.cuda.device(self.device):
            _lib.check(self.lib.bta_batch_potrf(self._h, _stream_ptr(self.device)), "bta_batch_potrf")
        for s, sp in zip(self.solvers, self.sparsities):
            s._factored = sp

    def status(self):
        return [s.status() for s in self.solvers]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self.lib.bta_batch_destroy(h)
            except Exception:
                pass
            self._h = None
