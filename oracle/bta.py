"""numpy/scipy restatement of the BTA block algorithms (CPU oracle).

Storage follows the reference's SerinvSolver buffers
(src/pyinla/solvers/structured_solver.py:36-60, scatter at :214-250):

    D[i] = Q[i*b:(i+1)*b, i*b:(i+1)*b]          i < n      (n, b, b)
    L[i] = Q[(i+1)*b:(i+2)*b, i*b:(i+1)*b]      i < n-1    (n-1, b, b)
    B[i] = Q[N-a:N, i*b:(i+1)*b]                i < n      (n, a, b)
    T    = Q[N-a:, N-a:]                                   (a, a)

with N = n*b + a.  The arithmetic itself lives in the external Serinv package
(imported at structured_solver.py:12-15, called in place at :83-93 (pobtaf /
pobtf), :112-140 (pobtas / pobts, trans "N" then "C"), :176-186 (pobtasi /
pobtsi)); Serinv is not vendored and not installed, so these functions restate
the published block recurrences of a block Cholesky factorisation of a
block-tridiagonal matrix with an arrowhead, its two triangular sweeps and the
Takahashi selected-inversion recurrence.  All functions work in place on float64
numpy arrays, as Serinv does.  ``B`` and ``T`` may be ``None`` (BT variant,
arrowhead size 0).
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla


def _chol(A: np.ndarray) -> np.ndarray:
    """Lower Cholesky factor with zero strict upper part (numpy semantics)."""
    return np.linalg.cholesky(A)


def _rtrsm_lt(X: np.ndarray, Lf: np.ndarray) -> np.ndarray:
    """Return X @ inv(Lf).T for lower-triangular Lf (right-side TRSM)."""
    # X Lf^T = Y  <=>  Lf X^T = Y^T
    return sla.solve_triangular(Lf, X.T, lower=True, check_finite=False).T


def pobtaf(D, L, B=None, T=None):
    """In-place BTA Cholesky: Q = LL^T, block lower-bidiagonal with an arrow row.

    Follows the call at structured_solver.py:83-88 (pobtaf(D, L, B, T)) and
    :90-93 (pobtf(D, L) when B is None).  Recurrence, for i = 0..n-1:
        D[i] <- chol(D[i])
        L[i] <- L[i] D[i]^-T ;  D[i+1] -= L[i] L[i]^T          (i < n-1)
        B[i] <- B[i] D[i]^-T ;  T -= B[i] B[i]^T
        B[i+1] -= B[i] L[i]^T                                   (i < n-1)
    then T <- chol(T).
    """
    n = D.shape[0]
    arrow = B is not None and T is not None and B.shape[1] > 0
    for i in range(n):
        D[i] = _chol(D[i])
        if i < n - 1:
            L[i] = _rtrsm_lt(L[i], D[i])
            D[i + 1] -= L[i] @ L[i].T
        if arrow:
            B[i] = _rtrsm_lt(B[i], D[i])
            T -= B[i] @ B[i].T
            if i < n - 1:
                B[i + 1] -= B[i] @ L[i].T
    if arrow:
        T[:] = _chol(T)


def pobtf(D, L):
    """BT variant of :func:`pobtaf` (structured_solver.py:90-93)."""
    pobtaf(D, L, None, None)


def logdet(D, T=None) -> float:
    """2 * (sum_i sum log diag D[i] + sum log diag T) (structured_solver.py:151-166).

    Raises ValueError on NaN, as the reference does at :163-164.
    """
    ld = 0.0
    for i in range(D.shape[0]):
        ld += float(np.sum(np.log(np.diagonal(D[i]))))
    if T is not None and T.shape[0] > 0:
        ld += float(np.sum(np.log(np.diagonal(T))))
    if np.isnan(ld):
        raise ValueError("Logdet is NaN. Check the input matrix.")
    return 2.0 * ld


def pobtas(D, L, B, T, rhs, trans: str = "N"):
    """In-place triangular sweep with the BTA factor (structured_solver.py:112-127).

    trans="N": solve  LL y = rhs  (forward);  trans="C"/"T": solve LL^T x = rhs.
    ``rhs`` has N rows and any number of columns (1-D allowed).
    """
    n, b = D.shape[0], D.shape[1]
    arrow = B is not None and T is not None and B.shape[1] > 0
    a = B.shape[1] if arrow else 0
    vec = rhs.ndim == 1
    R = rhs.reshape(rhs.shape[0], -1)
    nb = n * b
    if trans == "N":
        for i in range(n):
            r = R[i * b:(i + 1) * b]
            if i > 0:
                r -= L[i - 1] @ R[(i - 1) * b:i * b]
            r[:] = sla.solve_triangular(D[i], r, lower=True, check_finite=False)
        if arrow:
            ra = R[nb:nb + a]
            for i in range(n):
                ra -= B[i] @ R[i * b:(i + 1) * b]
            ra[:] = sla.solve_triangular(T, ra, lower=True, check_finite=False)
    elif trans in ("C", "T"):
        if arrow:
            ra = R[nb:nb + a]
            ra[:] = sla.solve_triangular(T, ra, lower=True, trans="T", check_finite=False)
        for i in range(n - 1, -1, -1):
            r = R[i * b:(i + 1) * b]
            if i < n - 1:
                r -= L[i].T @ R[(i + 1) * b:(i + 2) * b]
            if arrow:
                r -= B[i].T @ R[nb:nb + a]
            r[:] = sla.solve_triangular(D[i], r, lower=True, trans="T", check_finite=False)
    else:
        raise ValueError(f"unknown trans {trans!r}")
    if vec:
        rhs[:] = R[:, 0]
    return rhs


def pobts(D, L, rhs, trans: str = "N"):
    """BT variant of :func:`pobtas` (structured_solver.py:129-140)."""
    return pobtas(D, L, None, None, rhs, trans)


def pobtasi(D, L, B=None, T=None):
    """In-place selected inversion Sigma = Q^-1 on the BTA pattern.

    Follows the call at structured_solver.py:176-186.  With the factor blocks
    (D_i lower, L_i, B_i, T) and W_i = D_i^-1, G_i = L_i W_i, Ga_i = B_i W_i:
        Sigma_T      = T^-T T^-1
        for i = n-1 .. 0:
          Sigma_{n,i}   = -Sigma_{n,i+1} G_i - Sigma_T Ga_i   (first term only if i<n-1)
          Sigma_{i+1,i} = -Sigma_{i+1,i+1} G_i - Sigma_{n,i+1}^T Ga_i
          Sigma_{ii}    = W_i^T W_i - G_i^T Sigma_{i+1,i} - Ga_i^T Sigma_{n,i}
    Sigma_ii is written full symmetric into D[i], Sigma_{i+1,i} into L[i],
    Sigma_{n,i} into B[i] and Sigma_T into T.
    """
    n, b = D.shape[0], D.shape[1]
    arrow = B is not None and T is not None and B.shape[1] > 0
    if arrow:
        Tinv = sla.solve_triangular(T, np.eye(T.shape[0]), lower=True, check_finite=False)
        T[:] = Tinv.T @ Tinv
    for i in range(n - 1, -1, -1):
        W = sla.solve_triangular(D[i], np.eye(b), lower=True, check_finite=False)
        if i < n - 1:
            G = L[i] @ W
        if arrow:
            Ga = B[i] @ W
            Sn = -(T @ Ga)
            if i < n - 1:
                Sn -= B[i + 1] @ G
        if i < n - 1:
            Sl = -(D[i + 1] @ G)
            if arrow:
                Sl -= B[i + 1].T @ Ga
        Sd = W.T @ W
        if i < n - 1:
            Sd -= G.T @ Sl
        if arrow:
            Sd -= Ga.T @ Sn
        Sd = 0.5 * (Sd + Sd.T)
        if i < n - 1:
            L[i] = Sl
        if arrow:
            B[i] = Sn
        D[i] = Sd


def pobtsi(D, L):
    """BT variant of :func:`pobtasi` (structured_solver.py:182-186)."""
    pobtasi(D, L, None, None)


# ---------------------------------------------------------------- conversions

def bta_to_dense(D, L, B=None, T=None, lower_only: bool = False) -> np.ndarray:
    """Assemble the dense symmetric N x N matrix from BTA blocks."""
    n, b = D.shape[0], D.shape[1]
    a = 0 if B is None else B.shape[1]
    N = n * b + a
    Q = np.zeros((N, N))
    for i in range(n):
        Q[i * b:(i + 1) * b, i * b:(i + 1) * b] = D[i]
        if i < n - 1:
            Q[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] = L[i]
        if a:
            Q[n * b:, i * b:(i + 1) * b] = B[i]
    if a:
        Q[n * b:, n * b:] = T
    if lower_only:
        return np.tril(Q)
    return np.tril(Q) + np.tril(Q, -1).T


def dense_to_bta(Q: np.ndarray, n: int, b: int, a: int):
    """Slice a dense matrix into (D, L, B, T) copies (B, T are None if a == 0)."""
    D = np.stack([Q[i * b:(i + 1) * b, i * b:(i + 1) * b] for i in range(n)]).copy()
    L = np.stack([Q[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] for i in range(n - 1)]).copy() \
        if n > 1 else np.zeros((0, b, b))
    if a:
        B = np.stack([Q[n * b:, i * b:(i + 1) * b] for i in range(n)]).copy()
        T = Q[n * b:, n * b:].copy()
    else:
        B = T = None
    return D, L, B, T


def factor_to_dense(D, L, B=None, T=None) -> np.ndarray:
    """Dense lower-triangular factor from factored BTA blocks."""
    return bta_to_dense(np.tril(D), L, B, None if T is None else np.tril(T), lower_only=True)


# --------------------------------------------------------------- dense oracle

def dense_oracle(Q: np.ndarray, n: int, b: int, a: int):
    """Dense FP64 oracle in the reference DenseSolver pattern (dense_solver.py:34-68).

    Returns (factor blocks, logdet, Sigma blocks) computed from the dense
    Cholesky factor and the dense inverse, sliced to the BTA pattern.
    """
    Lf = np.linalg.cholesky(Q)
    ld = 2.0 * float(np.sum(np.log(np.diag(Lf))))
    Linv = sla.solve_triangular(Lf, np.eye(Q.shape[0]), lower=True)
    Sigma = Linv.T @ Linv
    return dense_to_bta(Lf, n, b, a), ld, dense_to_bta(Sigma, n, b, a)


# ----------------------------------------------------------- algorithmic work

def flops_potrf(n: int, b: int, a: int) -> float:
    """Algorithmic FLOPs of the BTA Cholesky (a = 0 gives the BT variant)."""
    return (n * b**3 / 3 + (n - 1) * 2 * b**3 + n * (a * b**2 + a**2 * b)
            + (n - 1) * 2 * a * b**2 + a**3 / 3)
