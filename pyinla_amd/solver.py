"""Drop-in replacement of pyINLA's SerinvSolver on MI355X.

Mirrors ``pyinla.solvers.structured_solver.SerinvSolver``
(src/pyinla/solvers/structured_solver.py:19-550): same constructor, same
methods (``cholesky``, ``solve``, ``logdet``, ``selected_inversion``,
``_structured_to_spmatrix``, ``get_solver_memory``), same in-place semantics,
same ``sparsity`` strings and the same errors (ValueError for an unknown
sparsity or a NaN log-determinant).  The block buffers are float64 torch
tensors on the GPU; the arithmetic runs in libpyinla_bta.so.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os
import time

import numpy as np
import torch

from . import _lib


class SolverConfig:
    """Stand-in for pyinla.configs.pyinla_config.SolverConfig (pyinla_config.py:10-15)."""

    def __init__(self, type: str = "serinv", min_processes: int = 1):
        self.type = type
        self.min_processes = min_processes


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream_ptr(device) -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _check_sparsity(sparsity: str) -> None:
    if sparsity not in ("bt", "bta"):
        raise ValueError(f"Unknown sparsity pattern: {sparsity}. Use 'bt' or 'bta'.")


class SerinvSolver:
    """BTA solver with SerinvSolver's interface, backed by HIP kernels.

    Parameters follow structured_solver.py:22-29.  ``arrowhead_blocksize`` may
    be 0 (BT matrices; the arrow buffers are then ``None``).  ``use_graph``
    replays each factorization / selected inversion as a cached hipGraph.
    ``reserve`` sizes the device workspace once, as the reference allocates
    its buffers once in ``__init__``: "selinv" (default; factor, solve and
    selected inversion), "solve" (no selected inversion) or "factor" (no
    solve either); an operation beyond the reservation raises.
    """

    def __init__(self, config=None, diagonal_blocksize: int = None, n_diag_blocks: int = None,
                 arrowhead_blocksize: int = 0, device=None, use_graph: bool = True, reserve: str = "selinv",
                 **kwargs) -> None:
        self.config = config if config is not None else SolverConfig()
        if diagonal_blocksize is None or n_diag_blocks is None:
            raise ValueError("diagonal_blocksize and n_diag_blocks are required")
        if not torch.cuda.is_available():
            raise RuntimeError("SerinvSolver (MI355X) needs a ROCm GPU; none is visible")
        self.lib = _lib.load()
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.diagonal_blocksize = int(diagonal_blocksize)
        self.arrowhead_blocksize = int(arrowhead_blocksize)
        self.n_diag_blocks = int(n_diag_blocks)
        n, b, a = self.n_diag_blocks, self.diagonal_blocksize, self.arrowhead_blocksize
        if a > 64:
            raise ValueError("arrowhead_blocksize > 64 is not supported")
        f64 = dict(dtype=torch.float64, device=self.device)
        self.A_diagonal_blocks = torch.empty((n, b, b), **f64)
        self.A_lower_diagonal_blocks = torch.empty((max(n - 1, 0), b, b), **f64)
        self.A_arrow_bottom_blocks = torch.empty((n, a, b), **f64) if a > 0 else None
        self.A_arrow_tip_block = torch.empty((a, a), **f64) if a > 0 else None
        if reserve not in _lib.RESERVE:
            raise ValueError(f"unknown reserve {reserve!r}: use 'factor', 'solve' or 'selinv'")
        flags = (_lib.WS_GRAPH if use_graph else 0) | _lib.RESERVE[reserve]
        ws = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.bta_ws_create(self.device.index or 0, n, b, a, flags, C.byref(ws)),
                       "bta_ws_create")
        self._ws = ws
        self._ld_out = torch.zeros(1, **f64)
        self._scatter_cache = {}
        self.total_bytes = 0
        self.t_cholesky = 0.0
        self.t_solve = 0.0
        self._factored = None  # sparsity of the factor currently held, or None
        # PYINLA_AMD_PANEL=auto|columns|flow selects the panel variant for every solver
        panel = os.environ.get("PYINLA_AMD_PANEL")
        if panel:
            self.set_panel(panel)

    def __del__(self):
        ws = getattr(self, "_ws", None)
        if ws is not None and ws.value:
            try:
                self.lib.bta_ws_destroy(ws)
            except Exception:
                pass
            self._ws = None

    # ------------------------------------------------------------ internals

    def _blocks(self, sparsity: str):
        a = self.arrowhead_blocksize if sparsity == "bta" else 0
        B = self.A_arrow_bottom_blocks if a > 0 else None
        T = self.A_arrow_tip_block if a > 0 else None
        return self.A_diagonal_blocks, self.A_lower_diagonal_blocks, B, T, a

    def factorize(self, sparsity: str = "bta") -> None:
        """Factor the blocks already held in the buffers (no scatter)."""
        _check_sparsity(sparsity)
        D, L, B, T, a = self._blocks(sparsity)
        n, b = self.n_diag_blocks, self.diagonal_blocksize
        with torch.cuda.device(self.device):
            _lib.check(self.lib.bta_potrf(self._ws, _ptr(D), _ptr(L), _ptr(B), _ptr(T), n, b, a,
                                          _stream_ptr(self.device)), "bta_potrf")
        self._factored = sparsity

    def set_panel(self, mode: str = "auto", helpers: int = 0, delta: int | None = None,
                  fold: str | None = None) -> None:
        """Panel factorization variant of ``factorize`` (no reference
        counterpart): "auto", "columns" (one launch per 64-wide tile column)
        or "flow" (one persistent dataflow launch per block); ``helpers`` sets
        the flow launch's helper workgroups (0 = automatic) and ``delta`` the
        order of their task lists (0 = simulated critical-path order, d >= 1 =
        dependency levels with far-tile updates grouped by d columns);
        ``fold`` whether the flow launch applies the block update itself
        ("panel"), leaves it to a grouped GEMM after it ("gemm") or picks by
        the block size ("auto")."""
        modes = {"auto": 0, "columns": 1, "flow": 2}
        if mode not in modes:
            raise ValueError(f"unknown panel mode {mode!r}")
        folds = {"auto": 0, "panel": 1, "gemm": 2}
        if fold is not None and fold not in folds:
            raise ValueError(f"unknown block-update placement {fold!r}")
        _lib.check(self.lib.bta_ws_set_option(self._ws, 3, modes[mode]), "set_option")
        _lib.check(self.lib.bta_ws_set_option(self._ws, 4, int(helpers)), "set_option")
        if delta is not None:
            _lib.check(self.lib.bta_ws_set_option(self._ws, 5, int(delta)), "set_option")
        if fold is not None:
            _lib.check(self.lib.bta_ws_set_option(self._ws, 10, folds[fold]), "set_option")

    def set_batch_lanes(self, lanes: int = 0) -> None:
        """Problems factored side by side when this solver is the first member
        of a ``BatchedFactorization`` built afterwards (no reference
        counterpart; 0 = automatic)."""
        _lib.check(self.lib.bta_ws_set_option(self._ws, 6, int(lanes)), "set_option")

    def set_selinv_blas(self, on: bool = True) -> None:
        """Selected inversion: the b x b x b products Sigma_{i+1,i} =
        -Sigma_{i+1,i+1} G_i as rocBLAS dgemm calls (default) or in the grouped
        FP64-MFMA kernel with everything else (no reference counterpart)."""
        _lib.check(self.lib.bta_ws_set_option(self._ws, 7, 1 if on else 0), "set_option")

    def status(self):
        """(info, logdet, err) of the last factorization (synchronises)."""
        info = C.c_int64()
        ld = C.c_double()
        err = C.c_int()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.bta_ws_status(self._ws, C.byref(info), C.byref(ld), C.byref(err),
                                              _stream_ptr(self.device)), "bta_ws_status")
        if err.value:
            raise _lib.BTALibraryError("in-kernel wait timed out during the factorization")
        return int(info.value), float(ld.value)

    def load_blocks(self, D, L, B=None, T=None) -> None:
        """Copy BTA blocks (numpy or torch) into the solver buffers."""
        def cp(dst, src):
            if dst is None or src is None:
                return
            if isinstance(src, np.ndarray):
                src = torch.from_numpy(np.ascontiguousarray(src))
            dst.copy_(src.to(dtype=torch.float64), non_blocking=False)
        cp(self.A_diagonal_blocks, D)
        cp(self.A_lower_diagonal_blocks, L)
        cp(self.A_arrow_bottom_blocks, B)
        cp(self.A_arrow_tip_block, T)

    # ------------------------------------------------------ Solver interface

    def cholesky(self, A, sparsity: str) -> None:
        """Scatter ``A`` into the blocks and factor in place (structured_solver.py:72-100)."""
        _check_sparsity(sparsity)
        self._spmatrix_to_structured(A, sparsity)
        tic = time.perf_counter()
        torch.cuda.synchronize(self.device)
        self.factorize(sparsity)
        info, _ = self.status()
        toc = time.perf_counter()
        self.t_cholesky += toc - tic
        if info != 0:
            raise np.linalg.LinAlgError(f"Matrix is not positive definite (first failing pivot at row {info - 1})")

    def logdet(self, sparsity: str) -> float:
        """2 sum log diag of the blocks held now (structured_solver.py:151-166):
        the diagonal blocks, plus the tip for "bta".  Like the reference it
        reads the current blocks, whichever factorization left them there."""
        _check_sparsity(sparsity)
        D, _, _, T, a = self._blocks(sparsity)
        n, b = self.n_diag_blocks, self.diagonal_blocksize
        with torch.cuda.device(self.device):
            _lib.check(self.lib.bta_logdet(self._ws, _ptr(D), _ptr(T), n, b, a, _ptr(self._ld_out),
                                           _stream_ptr(self.device)), "bta_logdet")
        ld = float(self._ld_out.item())
        if np.isnan(ld):
            raise ValueError("Logdet is NaN. Check the input matrix.")
        return ld

    def solve(self, rhs, sparsity: str):
        """Solve Q x = rhs with the held factor, in place (structured_solver.py:102-149).

        ``rhs`` is a float64 torch tensor on the solver's device of shape (N,)
        or (N, nrhs), overwritten and returned; a numpy array is solved on the
        device and written back in place.
        """
        _check_sparsity(sparsity)
        tic = time.perf_counter()
        host = None
        if isinstance(rhs, np.ndarray):
            host = rhs
            x = torch.from_numpy(np.ascontiguousarray(rhs, dtype=np.float64)).to(self.device)
        else:
            x = rhs
        if x.dtype != torch.float64 or x.device != self.device or not x.is_contiguous():
            raise ValueError("rhs must be a contiguous float64 tensor on the solver's device")
        X = x.reshape(x.shape[0], -1)
        D, L, B, T, a = self._blocks(sparsity)
        n, b = self.n_diag_blocks, self.diagonal_blocksize
        if X.shape[0] != n * b + a:
            raise ValueError(f"rhs has {X.shape[0]} rows, expected {n * b + a}")
        nrhs = X.shape[1]
        st = _stream_ptr(self.device)
        with torch.cuda.device(self.device):
            for trans in (0, 1):
                _lib.check(self.lib.bta_potrs(self._ws, _ptr(D), _ptr(L), _ptr(B), _ptr(T), n, b, a,
                                              C.c_void_p(X.data_ptr()), nrhs, nrhs, trans, st), "bta_potrs")
        torch.cuda.synchronize(self.device)
        self.t_solve += time.perf_counter() - tic
        if host is not None:
            host[...] = x.cpu().numpy().reshape(host.shape)
            return host
        return rhs

    def prepare_block_inverses(self, sparsity: str) -> None:
        """Enqueue the block inverses of the current factor that the next
        solve / selected inversion would otherwise build first (no reference
        counterpart)."""
        _check_sparsity(sparsity)
        D = self._blocks(sparsity)[0]
        with torch.cuda.device(self.device):
            _lib.check(self.lib.bta_block_inverses(self._ws, _ptr(D), self.n_diag_blocks, self.diagonal_blocksize,
                                                   _stream_ptr(self.device)), "bta_block_inverses")

    def selected_inversion(self, sparsity: str) -> None:
        """Overwrite the factor with the selected inverse (structured_solver.py:168-190).

        D[i] receives Sigma_ii (full symmetric), L[i] Sigma_{i+1,i}, B[i]
        Sigma_{n,i} and T Sigma_T.
        """
        _check_sparsity(sparsity)
        D, L, B, T, a = self._blocks(sparsity)
        n, b = self.n_diag_blocks, self.diagonal_blocksize
        with torch.cuda.device(self.device):
            _lib.check(self.lib.bta_potri(self._ws, _ptr(D), _ptr(L), _ptr(B), _ptr(T), n, b, a,
                                          _stream_ptr(self.device)), "bta_potri")
        self._factored = None

    def memory_bytes(self) -> int:
        """Device bytes this solver holds now: the block buffers, the
        workspace (its reservations and cached launch plans) and the cached
        scatter maps."""
        nbytes = sum(t.numel() * t.element_size() for t in (self.A_diagonal_blocks, self.A_lower_diagonal_blocks,
                                                            self.A_arrow_bottom_blocks, self.A_arrow_tip_block)
                     if t is not None)
        nbytes += self._ld_out.numel() * 8
        nbytes += int(self.lib.bta_ws_allocated(self._ws))
        nbytes += sum(e[0].numel() * 4 + e[1].numel() * 8 for e in self._scatter_cache.values())
        return nbytes

    def get_solver_memory(self) -> int:
        """Bytes held by the solver (``memory_bytes``), accumulated into
        ``total_bytes`` on every call as the reference does
        (structured_solver.py:540-550)."""
        self.total_bytes += self.memory_bytes()
        return self.total_bytes

    # ------------------------------------------------- sparse <-> blocks

    def _pattern_codes(self, rows, cols, sparsity: str, lower_only: bool):
        """Select the entries that the reference's scatter places in the blocks
        and return (entry indices, BTA codes).

        Placed (structured_solver.py:214-250): both triangles of the diagonal
        blocks, the sub-diagonal blocks (i+1, i), and for "bta" the arrow rows
        (n, i) and the tip.  lower_only keeps i >= j only (the gather of
        structured_solver.py:470-538 works on tril(A)).
        """
        from .st_model import bta_codes
        n, b = self.n_diag_blocks, self.diagonal_blocksize
        a = self.arrowhead_blocksize if sparsity == "bta" else 0
        nb_ = n * b
        rows = np.asarray(rows, dtype=np.int64)
        cols = np.asarray(cols, dtype=np.int64)
        st_r, st_c = rows < nb_, cols < nb_
        bi = np.where(st_r, rows // b, n)
        bj = np.where(st_c, cols // b, n)
        keepD = st_r & st_c & (bi == bj)
        keepL = st_r & st_c & (bi == bj + 1)
        keep = keepD | keepL
        if a > 0:
            keep |= (~st_r) & (rows < nb_ + a) & ((cols < nb_) | (cols < nb_ + a))
        if lower_only:
            keep &= rows >= cols
        idx = np.nonzero(keep)[0]
        r, c = rows[idx], cols[idx]
        # entries of the upper triangle of a diagonal block are stored as they
        # are (the blocks hold both triangles, as in the reference)
        upper = (r < c)
        codes = np.empty(len(idx), dtype=np.uint64)
        if (~upper).any():
            codes[~upper] = bta_codes(r[~upper], c[~upper], n, b, a)
        if upper.any():
            ru, cu = r[upper], c[upper]
            tip = (ru >= nb_)
            cd = np.empty(len(ru), dtype=np.uint64)
            if (~tip).any():  # inside a diagonal block: offset of (row, col) within D
                blk = ru[~tip] // b
                off = blk * b * b + (ru[~tip] % b) * b + (cu[~tip] % b)
                cd[~tip] = off.astype(np.uint64)
            if tip.any():
                off = (ru[tip] - nb_) * a + (cu[tip] - nb_)
                cd[tip] = (np.uint64(3) << np.uint64(60)) | off.astype(np.uint64)
            codes[upper] = cd
        return idx, codes

    @staticmethod
    def _pattern_key(rows, cols) -> bytes:
        """Digest of the full COO index arrays: two patterns share a scatter
        map only if every (row, col) in COO order agrees."""
        h = hashlib.blake2b(digest_size=20)
        h.update(np.ascontiguousarray(rows, dtype=np.int64).tobytes())
        h.update(np.ascontiguousarray(cols, dtype=np.int64).tobytes())
        return h.digest()

    def _spmatrix_to_structured(self, A, sparsity: str) -> None:
        """Zero the blocks and scatter the sparse matrix into them
        (structured_solver.py:192-250).  ``A`` is a scipy sparse matrix or a
        torch sparse tensor.  The scatter map is computed once per sparsity
        pattern and cached, as the reference's GPU path does (:258-365): the
        pattern and COO order must stay fixed between calls."""
        import scipy.sparse as sps
        _check_sparsity(sparsity)
        if isinstance(A, torch.Tensor):
            Ac = A.coalesce().cpu() if A.layout == torch.sparse_coo else A.to_sparse_coo().coalesce().cpu()
            ind = Ac.indices().numpy()
            rows, cols, data = ind[0], ind[1], Ac.values().numpy().astype(np.float64)
        else:
            coo = sps.coo_matrix(A)
            rows, cols, data = coo.row, coo.col, np.asarray(coo.data, dtype=np.float64)
        key = (sparsity, len(data), self._pattern_key(rows, cols))
        ent = self._scatter_cache.get(key)
        if ent is None:
            idx, codes = self._pattern_codes(rows, cols, sparsity, lower_only=False)
            ent = (torch.from_numpy(idx.astype(np.int32)).to(self.device),
                   torch.from_numpy(codes.view(np.int64)).to(self.device), len(idx))
            self._scatter_cache[key] = ent
        src_idx, codes, count = ent
        for t in (self.A_diagonal_blocks, self.A_lower_diagonal_blocks, self.A_arrow_bottom_blocks,
                  self.A_arrow_tip_block):
            if t is not None:
                t.zero_()
        vals = torch.from_numpy(np.ascontiguousarray(data)).to(self.device)
        D, L, B, T, a = self._blocks(sparsity)
        with torch.cuda.device(self.device):
            _lib.check(self.lib.bta_scatter_coo(_ptr(D), _ptr(L), _ptr(B), _ptr(T), _ptr(codes), _ptr(vals),
                                                _ptr(src_idx), count, 1.0, 0, _stream_ptr(self.device)),
                       "bta_scatter_coo")

    def _structured_to_spmatrix(self, A, sparsity: str):
        """Selected inverse on the lower pattern of ``A``, symmetrised
        (structured_solver.py:470-538).  Returns a scipy CSC matrix."""
        import scipy.sparse as sps
        _check_sparsity(sparsity)
        Bl = sps.coo_matrix(sps.tril(sps.csc_matrix(A)))
        idx, codes = self._pattern_codes(Bl.row, Bl.col, sparsity, lower_only=True)
        rows, cols = Bl.row[idx], Bl.col[idx]
        out = torch.empty(len(idx), dtype=torch.float64, device=self.device)
        codes_t = torch.from_numpy(codes.view(np.int64)).to(self.device)
        D, L, B, T, a = self._blocks(sparsity)
        with torch.cuda.device(self.device):
            _lib.check(self.lib.bta_gather_coo(_ptr(D), _ptr(L), _ptr(B), _ptr(T), _ptr(codes_t), _ptr(out),
                                               len(idx), _stream_ptr(self.device)), "bta_gather_coo")
        data = out.cpu().numpy()
        B_out = sps.coo_matrix((data, (rows, cols)), shape=Bl.shape).tocsc()
        return B_out + sps.tril(B_out, k=-1).T
