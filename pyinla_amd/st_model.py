"""Device-resident f(theta) and gradient of the Gaussian spatio-temporal model.

f follows PyINLA._evaluate_f for a Gaussian likelihood
(src/pyinla/core/pyinla.py:602-745 with :1076-1167), evaluated at x = 0,
eta = 0 as the reference does (pyinla.py:640-641):

    b      = e^{theta_o} A^T y                        (model.py:425-449)
    x*     = Q_c^-1 b,  Q_c = Q_p + e^{theta_o} A^T A  (model.py:384-423)
    f      = 1/2 ld(Q_c) - 1/2 x*^T Q_c x* - log p(theta)
             - (1/2 theta_o m - 1/2 e^{theta_o} y^T y) - 1/2 ld_BT(Q_p)

where ld_BT factors only the space-time (BT) part of Q_p (sparsity="bt" at
pyinla.py:1102-1103, so the fixed-effects tip is excluded).  x*^T Q_c x* is
evaluated as x*^T b, its value at x* = Q_c^-1 b.

Gradients:
  * finite differences exactly as PyINLA._objective_function
    (pyinla.py:512-600): 2d+1 evaluations at theta, theta +- eps e_i, run in
    reverse order, grad_i = (f[i+1] - f[d+i+1]) / (2 eps);
  * the adjoint: one forward keeping both factors, then the selected inverse
    of each in place and trace / quadratic contractions:
      df/dtheta_i = 1/2 tr(Sigma_c dQ_c) - 1/2 tr(Sigma_p^BT dQ_st)
                    - x*^T db + 1/2 x*^T dQ_c x* - d log p
                    - 1/2 m [i=o] + 1/2 e^{theta_o} y^T y [i=o].
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import scipy.sparse as sps
import torch

from . import _lib
from .coefficients import CoefficientModel
from .priors import priors_of
from .solver import SerinvSolver, _ptr, _stream_ptr

BUF_SHIFT = 60


def bta_codes(rows: np.ndarray, cols: np.ndarray, n: int, b: int, a: int) -> np.ndarray:
    """Map global (row, col) of the lower BTA pattern to (buffer << 60) | offset.

    Buffers 0..3 are D (n,b,b), L (n-1,b,b), B (n,a,b), T (a,a), as in the
    reference's scatter (structured_solver.py:214-250).  Entries outside the
    BTA pattern raise ValueError.
    """
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    nb_ = n * b
    code = np.empty(rows.shape, dtype=np.uint64)
    st_r = rows < nb_
    st_c = cols < nb_
    bi = np.where(st_r, rows // b, n)
    bj = np.where(st_c, cols // b, n)
    ri = np.where(st_r, rows % b, rows - nb_)
    cj = np.where(st_c, cols % b, cols - nb_)
    isD = st_r & st_c & (bi == bj)
    isL = st_r & st_c & (bi == bj + 1)
    isB = (~st_r) & st_c
    isT = (~st_r) & (~st_c)
    bad = ~(isD | isL | isB | isT)
    if bad.any():
        raise ValueError("matrix entries outside the lower BTA pattern")
    off = np.zeros(rows.shape, dtype=np.int64)
    off[isD] = bi[isD] * b * b + ri[isD] * b + cj[isD]
    off[isL] = bj[isL] * b * b + ri[isL] * b + cj[isL]
    off[isB] = bj[isB] * a * b + ri[isB] * b + cj[isB]
    off[isT] = ri[isT] * a + cj[isT]
    buf = np.zeros(rows.shape, dtype=np.uint64)
    buf[isL] = 1
    buf[isB] = 2
    buf[isT] = 3
    code[:] = (buf << np.uint64(BUF_SHIFT)) | off.astype(np.uint64)
    return code


class _ObsSet:
    """One observation block (A_v, y_v) with its own noise precision theta[idx]."""

    def __init__(self, A, y, idx, n, b, a, dev):
        A = sps.csr_matrix(A)
        AtA = sps.tril(sps.csr_matrix(A.T @ A)).tocoo()
        AtA.sum_duplicates()
        self.idx = idx
        self.code = torch.from_numpy(bta_codes(AtA.row, AtA.col, n, b, a).view(np.int64)).to(dev)
        self.val = torch.from_numpy(AtA.data.copy()).to(dev)
        w = np.where(AtA.row == AtA.col, 1.0, 2.0)
        self.wval = torch.from_numpy(AtA.data * w).to(dev)
        self.row = torch.from_numpy(AtA.row.astype(np.int32)).to(dev)
        self.col = torch.from_numpy(AtA.col.astype(np.int32)).to(dev)
        self.count = int(AtA.nnz)
        self.Aty = torch.from_numpy(np.asarray(A.T @ y, dtype=np.float64)).to(dev)
        self.yty = float(np.asarray(y) @ np.asarray(y))
        self.m = A.shape[0]
        self._A = A  # host CSR, uploaded on first use by device_csr()
        self._A_dev = None

    def device_csr(self, dev):
        """(rowptr, colidx, val) of A on the device (int32, int32, float64)."""
        if self._A_dev is None:
            A = self._A
            self._A_dev = (torch.from_numpy(A.indptr.astype(np.int32)).to(dev),
                           torch.from_numpy(A.indices.astype(np.int32)).to(dev),
                           torch.from_numpy(A.data.astype(np.float64)).to(dev))
        return self._A_dev


class STGaussianModel:
    """Gaussian (spatio-temporal or coregional) model with its two BTA solvers
    resident on the GPU.

    Holds Q_cond (BTA: n = nt, b = nv * ns, a = total fixed effects) and the
    space-time part of Q_prior (BT) in separate solvers so that both factors
    stay available for the adjoint.  ``coef`` maps theta to the coefficient
    table (see coefficients.py); nv = 1 is the single-field model, nv = 2 or 3
    the coregional model with fields interleaved per time block
    (coregional_model.py:754-786) and one noise precision per field.
    ``use_graph`` replays each solver operation from a hipGraph; it is off by
    default because graph replays of the two solvers do not overlap on the
    device (measured on MI355X at nt=200, ns=2000: the two factorizations
    take 537 ms with graphs and 365 ms launched directly on their streams).
    """

    def __init__(self, prob, coef=None, device=None, use_graph: bool = False):
        from .coefficients import CoregionalCoefficients
        self.prob = prob
        nv = getattr(prob, "nv", 1)
        if coef is None:
            coef = CoefficientModel("baseline") if nv == 1 else CoregionalCoefficients(nv)
        self.coef = coef
        self.nv = nv
        self.device = torch.device(device) if device is not None else torch.device("cuda", 0)
        self.lib = _lib.load()
        ns, nt = prob.ns, prob.nt
        nb_total = prob.nb
        self.ns = ns
        self.n, self.b, self.a = nt, nv * ns, nb_total
        dev = self.device
        self.nK = coef.nK
        self.d_q = getattr(coef, "d_q", getattr(coef, "d_st", 3))
        # spatial matrices on one union CSR pattern
        S = [sps.csr_matrix(s) for s in prob.spatial][: self.nK]
        while len(S) < self.nK:
            S.append(sps.csr_matrix((ns, ns)))
        union = abs(S[0])
        for s in S[1:]:
            union = union + abs(s)
        union = sps.csr_matrix(union)
        union.sum_duplicates()
        union.sort_indices()
        urow = np.repeat(np.arange(ns, dtype=np.int64), np.diff(union.indptr))
        ukey = urow * ns + union.indices.astype(np.int64)  # sorted (CSR, sorted indices)
        vals = []
        for s in S:
            c = sps.coo_matrix(s)
            c.sum_duplicates()
            key = c.row.astype(np.int64) * ns + c.col.astype(np.int64)
            pos = np.searchsorted(ukey, key)
            if not np.array_equal(ukey[np.minimum(pos, len(ukey) - 1)], key):
                raise ValueError("spatial matrix entry outside the union pattern")
            v = np.zeros(len(ukey))
            v[pos] = c.data
            vals.append(v)
        self.nnz = union.nnz
        self.rowptr = torch.from_numpy(union.indptr.astype(np.int32)).to(dev)
        self.colidx = torch.from_numpy(union.indices.astype(np.int32)).to(dev)
        self.vals = torch.from_numpy(np.concatenate(vals)).to(dev)
        # temporal tridiagonals
        nM = len(prob.temporal)
        Md = np.zeros((nM, nt))
        Mo = np.zeros((nM, max(nt - 1, 1)))
        for m, M in enumerate(prob.temporal):
            M = sps.csr_matrix(M)
            if nt > 2 and abs(sps.triu(M, 2)).sum() > 0:
                raise ValueError("temporal matrices must be tridiagonal")
            Md[m] = M.diagonal()
            if nt > 1:
                Mo[m, : nt - 1] = M.diagonal(-1)
        self.Mdiag_np = Md
        self.Moff_np = Mo[:, : nt - 1] if nt > 1 else Mo[:, :0]
        self.Mdiag = torch.from_numpy(Md).to(dev)
        self.Moff = torch.from_numpy(np.ascontiguousarray(Mo)).to(dev)
        self.nM = nM
        # observation blocks, each with its noise precision index in theta
        n, b, a = self.n, self.b, self.a
        noise = getattr(coef, "noise_index", lambda v: self.d_q + v)
        if getattr(prob, "A_list", None) is not None:
            idxs = [noise(v) for v in range(len(prob.A_list))]
            self.obs = [_ObsSet(A, y, i, n, b, a, dev) for A, y, i in zip(prob.A_list, prob.y_list, idxs)]
        else:
            self.obs = [_ObsSet(prob.A, prob.y, noise(0), n, b, a, dev)]
        self.priors = priors_of(prob)
        self.m_obs = sum(o.m for o in self.obs)
        # solvers: conditional (BTA) and prior space-time part (BT)
        self.cond = SerinvSolver(None, b, nt, a, device=dev, use_graph=use_graph)
        self.prior = SerinvSolver(None, b, nt, 0, device=dev, use_graph=use_graph)
        self.x = torch.zeros(nt * b + a, dtype=torch.float64, device=dev)
        self.rhs = torch.zeros_like(self.x)
        chunks = (ns + 63) // 64
        self.work = torch.zeros(max(nt * nv * chunks * 2 * nv * self.nK, 4096), dtype=torch.float64, device=dev)
        self.contract_out = torch.zeros(nt * nv * nv * self.nK * 2, dtype=torch.float64, device=dev)
        self.work_p = torch.zeros_like(self.work)
        self.contract_out_p = torch.zeros_like(self.contract_out)
        self.scalar = torch.zeros(4, dtype=torch.float64, device=dev)
        # the conditional and the prior chains are independent: each gets a
        # stream of its own (default priority) and they overlap on the GPU
        self.s_cond = torch.cuda.Stream(device=dev)
        self.s_prior = torch.cuda.Stream(device=dev)

    # ---------------------------------------------------------- assembly

    def _table(self, theta) -> np.ndarray:
        t = np.asarray(self.coef.table(theta), dtype=np.float64)
        return t.reshape(self.nv, self.nv, self.nM, self.nK)

    def _jacobian(self, theta) -> np.ndarray:
        J = np.asarray(self.coef.jacobian(theta), dtype=np.float64)
        return J.reshape(J.shape[0], self.nv, self.nv, self.nM, self.nK)

    def _assemble(self, solver: SerinvSolver, table: np.ndarray, with_arrow: bool) -> None:
        n, b, a = self.n, self.b, self.a
        coef = np.ascontiguousarray(table, dtype=np.float64).ravel()
        arrow_a = a if with_arrow else 0
        Bp = solver.A_arrow_bottom_blocks if with_arrow else None
        Tp = solver.A_arrow_tip_block if with_arrow else None
        _lib.check(self.lib.bta_assemble_kron(
            _ptr(solver.A_diagonal_blocks), _ptr(solver.A_lower_diagonal_blocks), _ptr(Bp), _ptr(Tp),
            n, b, arrow_a, self.nv, _ptr(self.rowptr), _ptr(self.colidx), _ptr(self.vals), self.nnz, self.nK,
            _ptr(self.Mdiag), _ptr(self.Moff), self.nM, coef.ctypes.data_as(C.c_void_p),
            float(self.prob.fixed_prec), _stream_ptr(self.device)), "bta_assemble_kron")

    def assemble_prior(self, theta, solver=None) -> None:
        self._assemble(solver or self.prior, self._table(theta), with_arrow=False)

    def assemble_cond(self, theta, solver=None) -> None:
        s = solver or self.cond
        self._assemble(s, self._table(theta), with_arrow=True)
        for o in self.obs:
            _lib.check(self.lib.bta_scatter_coo(
                _ptr(s.A_diagonal_blocks), _ptr(s.A_lower_diagonal_blocks), _ptr(s.A_arrow_bottom_blocks),
                _ptr(s.A_arrow_tip_block), _ptr(o.code), _ptr(o.val), None, o.count,
                float(np.exp(theta[o.idx])), 1, _stream_ptr(self.device)), "bta_scatter_coo")

    def _rhs(self, theta) -> None:
        self.rhs.zero_()
        for o in self.obs:
            self.rhs.add_(o.Aty, alpha=float(np.exp(theta[o.idx])))

    def _tail(self, theta, ld_c, ld_p, quad):
        logp = self.priors.log_prior(theta)
        lik = sum(0.5 * theta[o.idx] * o.m - 0.5 * np.exp(theta[o.idx]) * o.yty for o in self.obs)
        f = 0.5 * ld_c - 0.5 * quad - logp - lik - 0.5 * ld_p
        if np.isnan(f):
            raise ValueError("f(theta) is NaN. Check what is happening.")
        return dict(f=f, ld_c=ld_c, ld_p=ld_p, quad=quad, logp=logp, lik=lik)

    # -------------------------------------------------------------- f(theta)

    def _fork(self):
        cur = torch.cuda.current_stream(self.device)
        self.s_cond.wait_stream(cur)
        self.s_prior.wait_stream(cur)
        return cur

    def _join(self, cur):
        cur.wait_stream(self.s_cond)
        cur.wait_stream(self.s_prior)

    def forward(self, theta, prior_inverses: bool = False) -> dict:
        """Assemble, factor both matrices (on two streams), solve for x*;
        returns the pieces of f.  ``prior_inverses`` also builds the prior's
        block inverses (needed by its selected inversion) right behind its
        factorization, where they overlap the conditional solve."""
        theta = np.asarray(theta, dtype=np.float64)
        self._saved = None  # the factors about to be formed belong to no backward pass yet
        cur = self._fork()
        with torch.cuda.stream(self.s_prior):
            self.assemble_prior(theta)
            self.prior.factorize("bt")
            if prior_inverses:
                self.prior.prepare_block_inverses("bt")
        with torch.cuda.stream(self.s_cond):
            self.assemble_cond(theta)
            self.cond.factorize("bta")
            self._rhs(theta)
            self.x.copy_(self.rhs)
            self.cond.solve(self.x, "bta")
            info_c, ld_c = self.cond.status()
        with torch.cuda.stream(self.s_prior):
            info_p, ld_p = self.prior.status()
        self._join(cur)
        if info_c or info_p:
            raise np.linalg.LinAlgError(f"not positive definite (cond info {info_c}, prior info {info_p})")
        quad = float(torch.dot(self.x, self.rhs))
        return self._tail(theta, ld_c, ld_p, quad)

    def f(self, theta) -> float:
        return self.forward(theta)["f"]

    @staticmethod
    def _stencil(theta, eps):
        d = theta.size
        mat = np.repeat(theta[:, None], 2 * d + 1, axis=1)
        mat[:, 1:1 + d] += eps * np.eye(d)
        mat[:, d + 1:] -= eps * np.eye(d)
        return mat

    def fd_gradient(self, theta, eps: float = 1e-3):
        """Central differences as PyINLA._objective_function (pyinla.py:512-600)."""
        theta = np.asarray(theta, dtype=np.float64)
        d = theta.size
        mat = self._stencil(theta, eps)
        fv = np.zeros(2 * d + 1)
        for k in range(2 * d, -1, -1):
            fv[k] = self.f(mat[:, k])
        return fv[0], (fv[1:1 + d] - fv[d + 1:]) / (2 * eps)

    def fd_gradient_batched(self, theta, eps: float = 1e-3, wave: int | None = None):
        """The same FD gradient with the factorizations of up to ``wave``
        stencil points (default all 2d+1) in one batched, graph-replayed launch
        sequence; one extra solver pair per stencil point of a wave."""
        theta = np.asarray(theta, dtype=np.float64)
        d = theta.size
        mat = self._stencil(theta, eps)
        fv = self.f_points(mat, wave=wave, order=list(range(2 * d, -1, -1)))
        return fv[0], (fv[1:1 + d] - fv[d + 1:]) / (2 * eps)

    def f_points(self, mat, wave: int | None = None, order=None):
        """f at every column of ``mat`` (dim theta x npts), the factorizations
        of up to ``wave`` points (default all) in one batched launch sequence;
        ``order`` is the evaluation order of the columns (default: as given)."""
        from .batch import BatchedFactorization
        mat = np.asarray(mat, dtype=np.float64)
        npts = mat.shape[1]
        if wave is None:
            # as many extra solver pairs (factor blocks of both matrices, the
            # conditional's block inverses, with margin) as fit in 70 % of the
            # free device memory
            free, _ = torch.cuda.mem_get_info(self.device)
            per_pair = 8 * 6 * self.n * self.b * self.b
            wave = 1 + int(0.7 * free // per_pair)
        wave = max(1, min(wave, npts))
        pool = getattr(self, "_fd_pool", None)
        if pool is None or len(pool[0]) != wave:
            self._fd_pool = None
            n, b, a = self.n, self.b, self.a
            conds = [self.cond] + [SerinvSolver(None, b, n, a, device=self.device, reserve="solve")
                                   for _ in range(wave - 1)]
            priors = [self.prior] + [SerinvSolver(None, b, n, 0, device=self.device, reserve="factor")
                                     for _ in range(wave - 1)]
            batch = BatchedFactorization(conds + priors, ["bta"] * wave + ["bt"] * wave)
            self._fd_pool = (conds, priors, batch)
        conds, priors, batch = self._fd_pool
        self._saved = None  # the pool's first pair is the model's own solvers
        fv = np.zeros(npts)
        order = list(range(npts)) if order is None else list(order)
        for w0 in range(0, npts, wave):
            pts = order[w0:w0 + wave]
            for j, k in enumerate(pts):
                self.assemble_cond(mat[:, k], conds[j])
                self.assemble_prior(mat[:, k], priors[j])
            batch.factorize()
            for j, k in enumerate(pts):
                th = mat[:, k]
                self._rhs(th)
                self.x.copy_(self.rhs)
                conds[j].solve(self.x, "bta")
                info_c, ld_c = conds[j].status()
                info_p, ld_p = priors[j].status()
                if info_c or info_p:
                    raise np.linalg.LinAlgError("not positive definite in the FD batch")
                quad = float(torch.dot(self.x, self.rhs))
                fv[k] = self._tail(th, ld_c, ld_p, quad)["f"]
        return fv

    @staticmethod
    def _hessian_stencil(theta, eps):
        """Points of PyINLA._evaluate_hessian_f (pyinla.py:767-910): theta,
        theta +- eps e_i, and theta +- eps e_i +- eps e_j for i < j."""
        d = theta.size
        E = eps * np.eye(d)
        pts = [theta.copy()]
        for i in range(d):
            pts += [theta + E[i], theta - E[i]]
        for i in range(d):
            for j in range(i + 1, d):
                pts += [theta + E[i] + E[j], theta + E[i] - E[j], theta - E[i] + E[j], theta - E[i] - E[j]]
        return np.stack(pts, axis=1)

    def hessian_fd(self, theta, eps: float = 5e-3, batched: bool = True, wave: int | None = None):
        """Central-difference Hessian of f as PyINLA._evaluate_hessian_f
        computes it (pyinla.py:767-910, default eps_hessian_f = 5e-3,
        pyinla_config.py:52): 1 + 2d + 2d(d-1) evaluations of f, batched (the
        factorizations of ``wave`` points per launch sequence) or one by one."""
        theta = np.asarray(theta, dtype=np.float64)
        d = theta.size
        mat = self._hessian_stencil(theta, eps)
        if batched:
            fv = self.f_points(mat, wave=wave)
        else:
            fv = np.array([self.f(mat[:, k]) for k in range(mat.shape[1])])
        f0 = fv[0]
        H = np.zeros((d, d))
        for i in range(d):
            H[i, i] = (fv[1 + 2 * i] - 2.0 * f0 + fv[2 + 2 * i]) / eps ** 2
        q = 1 + 2 * d
        for i in range(d):
            for j in range(i + 1, d):
                fpp, fpm, fmp, fmm = fv[q:q + 4]
                H[i, j] = H[j, i] = (fpp - fpm - fmp + fmm) / (4.0 * eps * eps)
                q += 4
        return H

    def hessian_adjoint(self, theta, eps: float = 5e-3):
        """Hessian of f from central differences of the adjoint gradient:
        2d evaluations of value_and_grad instead of 1 + 2d + 2d(d-1) of f,
        symmetrised (no reference counterpart)."""
        theta = np.asarray(theta, dtype=np.float64)
        d = theta.size
        H = np.zeros((d, d))
        for i in range(d):
            e = np.zeros(d)
            e[i] = eps
            _, gp = self.value_and_grad(theta + e)
            _, gm = self.value_and_grad(theta - e)
            H[:, i] = (gp - gm) / (2.0 * eps)
        return 0.5 * (H + H.T)

    def cov_theta(self, theta, eps: float = 5e-3, method: str = "fd"):
        """Covariance of theta as the inverse of the Hessian of f
        (PyINLA.compute_covariance_hp, pyinla.py:747-765); ``method`` "fd"
        (the reference's stencil) or "adjoint" (differences of gradients)."""
        H = self.hessian_fd(theta, eps) if method == "fd" else self.hessian_adjoint(theta, eps)
        return np.linalg.inv(H)

    # ---------------------------------------------------------- adjoint

    def _contract_launch(self, D, L, x, mode, work=None, out=None):
        """Enqueue the Kronecker contraction on the current stream; returns the
        device output (n, nv, nv, nK, 2)."""
        n, b, nv, nK = self.n, self.b, self.nv, self.nK
        work = self.work if work is None else work
        out_t = self.contract_out if out is None else out
        _lib.check(self.lib.bta_kron_contract(
            _ptr(D), _ptr(L), _ptr(x), n, b, nv, _ptr(self.rowptr), _ptr(self.colidx), _ptr(self.vals),
            self.nnz, nK, mode, _ptr(work), _ptr(out_t), _stream_ptr(self.device)),
            "bta_kron_contract")
        return out_t[: n * nv * nv * nK * 2].reshape(n, nv, nv, nK, 2)

    def _contract_finish(self, out_dev):
        """Z[v][w][m][k] = sum_t Mdiag[m][t] PD[t][v][w][k] + 2 sum_t Moff[m][t] PL[t][v][w][k]."""
        n = self.n
        out = out_dev.cpu().numpy()
        PD, PL = out[..., 0], out[..., 1]
        Z = np.einsum("mt,tvwk->vwmk", self.Mdiag_np, PD)
        if n > 1:
            Z = Z + 2.0 * np.einsum("mt,tvwk->vwmk", self.Moff_np, PL[: n - 1])
        return Z

    def _contract(self, D, L, x, mode, work=None, out=None):
        return self._contract_finish(self._contract_launch(D, L, x, mode, work, out))

    def _bilinear(self, o: _ObsSet, mode, solver=None, x=None):
        s = solver
        nwork = min(1024, self.work.numel())
        _lib.check(self.lib.bta_bilinear_coo(
            _ptr(s.A_diagonal_blocks) if s else None, _ptr(s.A_lower_diagonal_blocks) if s else None,
            _ptr(s.A_arrow_bottom_blocks) if s else None, _ptr(s.A_arrow_tip_block) if s else None,
            _ptr(o.code), _ptr(o.row), _ptr(o.col), _ptr(x), _ptr(o.wval),
            o.count, mode, _ptr(self.work), nwork, _ptr(self.scalar), _stream_ptr(self.device)),
            "bta_bilinear_coo")
        return float(self.scalar[0])

    def forward_keep(self, theta) -> dict:
        """The adjoint's forward pass: f(theta) with both factors kept in HBM
        for the backward pass (and the prior's block inverses built behind
        its factorization, where they overlap the conditional solve).
        Returns the pieces of f; ``backward_grad`` consumes the saved state."""
        theta = np.asarray(theta, dtype=np.float64)
        parts = self.forward(theta, prior_inverses=True)
        self._epoch = getattr(self, "_epoch", 0) + 1
        self._saved = (self._epoch, theta.copy())
        return parts

    def backward_grad(self, token=None) -> np.ndarray:
        """The adjoint's backward pass over the factors ``forward_keep`` left:
        the selected inversion of each factor in place (pobtasi), the trace
        and quadratic-form contractions, and the prior / likelihood terms.
        ``token`` (the epoch forward_keep returned through ``saved_token``)
        guards against factors overwritten by a later evaluation: if they no
        longer belong to that forward, the forward is run again first."""
        saved = getattr(self, "_saved", None)
        if saved is None or (token is not None and token[0] != saved[0]):
            if token is None:
                raise RuntimeError("backward_grad needs a preceding forward_keep")
            self.forward_keep(token[1])
            saved = self._saved
        theta = saved[1]
        self._saved = None  # the selected inversions below overwrite the factors
        # the prior's selected inversion does not need x*: queued before the
        # host waits for the x* terms below
        cur = self._fork()
        with torch.cuda.stream(self.s_prior):
            self.prior.selected_inversion("bt")
            op = self._contract_launch(self.prior.A_diagonal_blocks, self.prior.A_lower_diagonal_blocks, None, 0,
                                       work=self.work_p, out=self.contract_out_p)
        # terms that need x* only
        Zq = self._contract(None, None, self.x, 1)
        xAtAx = [self._bilinear(o, 1, x=self.x) for o in self.obs]
        xAty = [float(torch.dot(self.x, o.Aty)) for o in self.obs]
        self.s_cond.wait_stream(cur)
        with torch.cuda.stream(self.s_cond):
            self.cond.selected_inversion("bta")
            oc = self._contract_launch(self.cond.A_diagonal_blocks, self.cond.A_lower_diagonal_blocks, None, 0)
        self._join(cur)
        Zp = self._contract_finish(op)
        Zc = self._contract_finish(oc)
        trAtA = [self._bilinear(o, 0, solver=self.cond) for o in self.obs]
        J = self._jacobian(theta)
        g = np.zeros(theta.size)
        Zsum = Zc - Zp + Zq
        for i in range(J.shape[0]):
            g[i] = 0.5 * np.sum(J[i] * Zsum)
        for o, tr, q, xb in zip(self.obs, trAtA, xAtAx, xAty):
            eo = float(np.exp(theta[o.idx]))
            g[o.idx] += 0.5 * eo * tr - eo * xb + 0.5 * eo * q - 0.5 * o.m + 0.5 * eo * o.yty
        g -= self.priors.grad(theta)
        return g

    @property
    def saved_token(self):
        """(epoch, theta) of the factors the last forward_keep left in HBM."""
        return None if getattr(self, "_saved", None) is None else (self._saved[0], self._saved[1].copy())

    def autotune(self, theta=None, reps: int = 2, cache_path: str | None = None) -> dict:
        """Tune the launch options of both solvers (pyinla_amd.autotune) on
        Q_prior(theta) and Q_cond(theta); the choices stay set on the
        workspaces.  Returns {"prior": choice, "cond": choice}."""
        from . import autotune as at
        theta = np.asarray(self.prob.theta if theta is None else theta, dtype=np.float64)
        self._saved = None
        self.assemble_prior(theta)
        cp = at.autotune(self.prior, "bt", reps=reps, cache_path=cache_path)
        self.assemble_cond(theta)
        cc = at.autotune(self.cond, "bta", reps=reps, cache_path=cache_path)
        torch.cuda.synchronize(self.device)
        return {"prior": cp, "cond": cc}

    def value_and_grad(self, theta):
        """f(theta) and its adjoint gradient: forward_keep then backward_grad."""
        parts = self.forward_keep(theta)
        return parts["f"], self.backward_grad()

    def marginal_variances(self, theta, observations: bool = True):
        """Marginal variances of the latent parameters and of the
        observations at theta, as PyINLA.get_marginal_variances_latent_parameters
        and get_marginal_variances_observations compute them for a Gaussian
        likelihood (pyinla.py:912-1016): Q_cond(theta) is factored and inverted
        on its BTA pattern; diag(Sigma) is read off the diagonal blocks (latent
        order of the solver: time blocks, then the fixed effects), and
        diag(A Sigma A^T) is formed per observation row on the device
        (bta_quadform_rows) instead of the reference's sparse triple product.
        Returns (latent variances, [observation variances per observation
        set]) as device tensors.  Overwrites the conditional solver's factor."""
        theta = np.asarray(theta, dtype=np.float64)
        self._saved = None
        s = self.cond
        self.assemble_cond(theta)
        s.factorize("bta")
        info, _ = s.status()
        if info:
            raise np.linalg.LinAlgError(f"not positive definite (cond info {info})")
        s.selected_inversion("bta")
        n, b, a = self.n, self.b, self.a
        parts = [torch.diagonal(s.A_diagonal_blocks, dim1=1, dim2=2).reshape(-1)]
        if a:
            parts.append(torch.diagonal(s.A_arrow_tip_block))
        latent = torch.cat(parts)
        obs = []
        if observations:
            for o in self.obs:
                rp, ci, v = o.device_csr(self.device)
                out = torch.empty(o.m, dtype=torch.float64, device=self.device)
                _lib.check(self.lib.bta_quadform_rows(
                    _ptr(s.A_diagonal_blocks), _ptr(s.A_lower_diagonal_blocks), _ptr(s.A_arrow_bottom_blocks),
                    _ptr(s.A_arrow_tip_block), n, b, a, _ptr(rp), _ptr(ci), _ptr(v), o.m, _ptr(out),
                    _stream_ptr(self.device)), "bta_quadform_rows")
                obs.append(out)
        return latent, obs


class AdjointF(torch.autograd.Function):
    """f(theta) as a differentiable torch scalar.

    forward: assemble and factor Q_prior and Q_cond, solve for x*, evaluate
    f; the factors stay in HBM.  backward: the selected inversion of both
    factors in place (pobtasi) and the trace / quadratic contractions give
    grad f; a forward that is never differentiated costs only the forward."""

    @staticmethod
    def forward(ctx, theta: torch.Tensor, model: STGaussianModel):
        parts = model.forward_keep(theta.detach().cpu().numpy())
        ctx.model = model
        ctx.token = model.saved_token
        ctx.theta_dtype = theta.dtype
        return torch.tensor(parts["f"], dtype=theta.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        g = ctx.model.backward_grad(ctx.token)
        return grad_out * torch.as_tensor(g, dtype=ctx.theta_dtype), None
