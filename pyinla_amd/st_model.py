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
from .coefficients import CoefficientModel, gaussian_log_prior, gaussian_log_prior_grad
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


class STGaussianModel:
    """Gaussian spatio-temporal model with its two BTA solvers resident on the GPU.

    Holds Q_cond (BTA: b = ns, n = nt, a = nb) and Q_prior's space-time part
    (BT) in separate solvers so that both factors stay available for the
    adjoint.  ``coef`` selects the coefficient table (see coefficients.py).
    """

    def __init__(self, prob, coef: CoefficientModel | None = None, device=None, use_graph: bool = True):
        self.prob = prob
        self.coef = coef if coef is not None else CoefficientModel("baseline")
        self.device = torch.device(device) if device is not None else torch.device("cuda", 0)
        self.lib = _lib.load()
        ns, nt, nb = prob.ns, prob.nt, prob.nb
        self.n, self.b, self.a = nt, ns, nb
        dev = self.device
        # spatial matrices on one union CSR pattern
        S = [sps.csr_matrix(s) for s in prob.spatial][: self.coef.nK]
        while len(S) < self.coef.nK:
            S.append(sps.csr_matrix((ns, ns)))
        union = abs(S[0])
        for s in S[1:]:
            union = union + abs(s)
        union = sps.csr_matrix(union)
        union.sort_indices()
        vals = []
        for s in S:
            s2 = s.tocsr().copy()
            s2.sort_indices()
            pat = union.copy()
            pat.data[:] = 0.0
            v = (pat + s2).tocsr()
            v.sort_indices()
            # same pattern as union (union contains every entry of s2)
            assert v.nnz == union.nnz
            vals.append(v.data.copy())
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
        self.Mdiag_np, self.Moff_np = Md, Mo[:, : max(nt - 1, 0)] if nt > 1 else Mo[:, :0]
        self.Mdiag = torch.from_numpy(Md).to(dev)
        self.Moff = torch.from_numpy(np.ascontiguousarray(Mo)).to(dev)
        self.nM = nM
        # A^T A on the lower BTA pattern, with symmetry weights
        A = sps.csr_matrix(prob.A)
        AtA = sps.tril(sps.csr_matrix(A.T @ A)).tocoo()
        AtA.sum_duplicates()
        self.ata_code = torch.from_numpy(bta_codes(AtA.row, AtA.col, nt, ns, nb).view(np.int64)).to(dev)
        self.ata_val = torch.from_numpy(AtA.data.copy()).to(dev)
        w = np.where(AtA.row == AtA.col, 1.0, 2.0)
        self.ata_wval = torch.from_numpy(AtA.data * w).to(dev)
        self.ata_row = torch.from_numpy(AtA.row.astype(np.int32)).to(dev)
        self.ata_col = torch.from_numpy(AtA.col.astype(np.int32)).to(dev)
        self.ata_count = int(AtA.nnz)
        self.Aty = torch.from_numpy(np.asarray(A.T @ prob.y, dtype=np.float64)).to(dev)
        self.yty = float(prob.y @ prob.y)
        self.m_obs = A.shape[0]
        # solvers: conditional (BTA) and prior space-time part (BT)
        self.cond = SerinvSolver(None, ns, nt, nb, device=dev, use_graph=use_graph)
        self.prior = SerinvSolver(None, ns, nt, 0, device=dev, use_graph=use_graph)
        self.x = torch.zeros(nt * ns + nb, dtype=torch.float64, device=dev)
        self.rhs = torch.zeros_like(self.x)
        chunks = (ns + 63) // 64
        self.work = torch.zeros(max(nt * chunks * 8, 4096), dtype=torch.float64, device=dev)
        self.contract_out = torch.zeros(nt * 8, dtype=torch.float64, device=dev)
        self.scalar = torch.zeros(4, dtype=torch.float64, device=dev)
        self.timers = {}

    # ---------------------------------------------------------- assembly

    def _assemble(self, solver: SerinvSolver, table: np.ndarray, with_arrow: bool) -> None:
        n, b, a = self.n, self.b, self.a
        coef = np.ascontiguousarray(table, dtype=np.float64).ravel()
        arrow_a = a if with_arrow else 0
        Bp = solver.A_arrow_bottom_blocks if with_arrow else None
        Tp = solver.A_arrow_tip_block if with_arrow else None
        _lib.check(self.lib.bta_assemble_kron(
            _ptr(solver.A_diagonal_blocks), _ptr(solver.A_lower_diagonal_blocks), _ptr(Bp), _ptr(Tp),
            n, b, arrow_a, _ptr(self.rowptr), _ptr(self.colidx), _ptr(self.vals), self.nnz, self.coef.nK,
            _ptr(self.Mdiag), _ptr(self.Moff), self.nM, coef.ctypes.data_as(C.c_void_p),
            float(self.prob.fixed_prec), _stream_ptr(self.device)), "bta_assemble_kron")

    def assemble_prior(self, theta) -> None:
        self._assemble(self.prior, self.coef.table(theta), with_arrow=False)

    def assemble_cond(self, theta) -> None:
        s = self.cond
        self._assemble(s, self.coef.table(theta), with_arrow=True)
        _lib.check(self.lib.bta_scatter_coo(
            _ptr(s.A_diagonal_blocks), _ptr(s.A_lower_diagonal_blocks), _ptr(s.A_arrow_bottom_blocks),
            _ptr(s.A_arrow_tip_block), _ptr(self.ata_code), _ptr(self.ata_val), None, self.ata_count,
            float(np.exp(theta[3])), 1, _stream_ptr(self.device)), "bta_scatter_coo")

    # -------------------------------------------------------------- f(theta)

    def forward(self, theta) -> dict:
        """Assemble, factor both matrices, solve for x*; returns the pieces of f."""
        theta = np.asarray(theta, dtype=np.float64)
        self.assemble_cond(theta)
        self.cond.factorize("bta")
        self.assemble_prior(theta)
        self.prior.factorize("bt")
        eo = float(np.exp(theta[3]))
        torch.mul(self.Aty, eo, out=self.rhs)
        self.x.copy_(self.rhs)
        self.cond.solve(self.x, "bta")
        info_c, ld_c = self.cond.status()
        info_p, ld_p = self.prior.status()
        if info_c or info_p:
            raise np.linalg.LinAlgError(f"not positive definite (cond info {info_c}, prior info {info_p})")
        quad = float(torch.dot(self.x, self.rhs))
        logp = gaussian_log_prior(theta, self.prob.prior_mean, self.prob.prior_prec)
        lik = 0.5 * theta[3] * self.m_obs - 0.5 * eo * self.yty
        f = 0.5 * ld_c - 0.5 * quad - logp - lik - 0.5 * ld_p
        if np.isnan(f):
            raise ValueError("f(theta) is NaN. Check what is happening.")
        return dict(f=f, ld_c=ld_c, ld_p=ld_p, quad=quad, logp=logp, lik=lik)

    def f(self, theta) -> float:
        return self.forward(theta)["f"]

    def fd_gradient(self, theta, eps: float = 1e-3):
        """Central differences as PyINLA._objective_function (pyinla.py:512-600)."""
        theta = np.asarray(theta, dtype=np.float64)
        d = theta.size
        mat = np.repeat(theta[:, None], 2 * d + 1, axis=1)
        mat[:, 1:1 + d] += eps * np.eye(d)
        mat[:, d + 1:] -= eps * np.eye(d)
        fv = np.zeros(2 * d + 1)
        for k in range(2 * d, -1, -1):
            fv[k] = self.f(mat[:, k])
        grad = (fv[1:1 + d] - fv[d + 1:]) / (2 * eps)
        return fv[0], grad

    # ---------------------------------------------------------- adjoint

    def _contract(self, D, L, x, mode):
        n, b = self.n, self.b
        nK = self.coef.nK
        _lib.check(self.lib.bta_kron_contract(
            _ptr(D), _ptr(L), _ptr(x), n, b, _ptr(self.rowptr), _ptr(self.colidx), _ptr(self.vals),
            self.nnz, nK, mode, _ptr(self.work), _ptr(self.contract_out), _stream_ptr(self.device)),
            "bta_kron_contract")
        out = self.contract_out[: n * 2 * nK].reshape(n, nK, 2).cpu().numpy()
        PD, PL = out[:, :, 0], out[:, :, 1]
        # Z[m][k] = sum_t Mdiag[m][t] PD[t][k] + 2 sum_t Moff[m][t] PL[t][k]
        Z = self.Mdiag_np @ PD
        if n > 1:
            Z = Z + 2.0 * (self.Moff_np @ PL[: n - 1])
        return Z

    def _bilinear(self, mode, solver=None, x=None):
        s = solver
        nwork = min(1024, self.work.numel())
        _lib.check(self.lib.bta_bilinear_coo(
            _ptr(s.A_diagonal_blocks) if s else None, _ptr(s.A_lower_diagonal_blocks) if s else None,
            _ptr(s.A_arrow_bottom_blocks) if s else None, _ptr(s.A_arrow_tip_block) if s else None,
            _ptr(self.ata_code), _ptr(self.ata_row), _ptr(self.ata_col), _ptr(x), _ptr(self.ata_wval),
            self.ata_count, mode, _ptr(self.work), nwork, _ptr(self.scalar), _stream_ptr(self.device)),
            "bta_bilinear_coo")
        return float(self.scalar[0])

    def value_and_grad(self, theta):
        """f(theta) and its adjoint gradient (one forward + selected inversions)."""
        theta = np.asarray(theta, dtype=np.float64)
        parts = self.forward(theta)
        eo = float(np.exp(theta[3]))
        # quadratic terms need x*, taken before the factors are overwritten
        Zq = self._contract(None, None, self.x, 1)
        xAtAx = self._bilinear(1, x=self.x)
        self.cond.selected_inversion("bta")
        self.prior.selected_inversion("bt")
        Zc = self._contract(self.cond.A_diagonal_blocks, self.cond.A_lower_diagonal_blocks, None, 0)
        Zp = self._contract(self.prior.A_diagonal_blocks, self.prior.A_lower_diagonal_blocks, None, 0)
        trAtA = self._bilinear(0, solver=self.cond)
        J = self.coef.jacobian(theta)  # (d_st, nM, nK)
        g = np.zeros(theta.size)
        dst = self.coef.d_st
        for i in range(dst):
            g[i] = 0.5 * np.sum(J[i] * Zc) - 0.5 * np.sum(J[i] * Zp) + 0.5 * np.sum(J[i] * Zq)
        g[3] = 0.5 * eo * trAtA - parts["quad"] + 0.5 * eo * xAtAx - 0.5 * self.m_obs + 0.5 * eo * self.yty
        g -= gaussian_log_prior_grad(theta, self.prob.prior_mean, self.prob.prior_prec)
        return parts["f"], g


class AdjointF(torch.autograd.Function):
    """f(theta) as a differentiable torch scalar; backward uses the adjoint."""

    @staticmethod
    def forward(ctx, theta: torch.Tensor, model: STGaussianModel):
        f, g = model.value_and_grad(theta.detach().cpu().numpy())
        ctx.save_for_backward(torch.as_tensor(g, dtype=theta.dtype))
        return torch.tensor(f, dtype=theta.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None
