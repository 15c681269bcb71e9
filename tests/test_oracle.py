"""Pin the CPU oracle (oracle/bta.py) against facts that do not come from it.

The reference's own tests hold no golden vectors for this path (its tests/
directory is empty and its example data are git-LFS pointers), so each oracle
function is pinned by:
  - Cholesky uniqueness on a known factor (the factor must be reproduced);
  - dense LAPACK (numpy cholesky / inv / slogdet) on the same matrix;
  - closed forms: logdet tridiag(-1, 2, -1)_n = ln(n+1); Kronecker identities
    logdet(M (x) S) = ns logdet M + nt logdet S, (M (x) S)^-1 = M^-1 (x) S^-1;
  - the Schur-complement identity for the arrowhead;
and the planted-mistake tests at the end show that plausible errors (a
transposed operand, a dropped term, a wrong sign) fail these checks.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import bta as ob
from pyinla_amd.workloads import known_factor_bta


def _random_bta(n, b, a, seed):
    kf = known_factor_bta(n, b, a, seed)
    return kf, ob.bta_to_dense(kf.Q_D, kf.Q_L, kf.Q_B if a else None, kf.Q_T if a else None)


def test_known_factor_plain_case_full_size():
    """n=50, b=92, a=4 (seed 0): pobtaf must return the generating factor."""
    kf = known_factor_bta()
    D, L, B, T = kf.Q_D.copy(), kf.Q_L.copy(), kf.Q_B.copy(), kf.Q_T.copy()
    ob.pobtaf(D, L, B, T)
    for got, exp in ((D, kf.Lf_D), (L, kf.Lf_L), (B, kf.Lf_B), (T, kf.Lf_T)):
        assert np.abs(got - exp).max() <= 1e-12 * np.abs(exp).max()
    ld_exact = 2 * (np.log(np.diagonal(kf.Lf_D, axis1=1, axis2=2)).sum() + np.log(np.diag(kf.Lf_T)).sum())
    assert abs(ob.logdet(D, T) - ld_exact) <= 1e-13 * abs(ld_exact)


@pytest.mark.parametrize("n,b,a", [(1, 5, 0), (1, 5, 3), (2, 7, 0), (4, 9, 2), (6, 3, 1)])
def test_pobtaf_matches_dense_lapack(n, b, a):
    kf, Q = _random_bta(n, b, a, seed=n * 10 + b + a)
    D, L = kf.Q_D.copy(), kf.Q_L.copy()
    B = kf.Q_B.copy() if a else None
    T = kf.Q_T.copy() if a else None
    ob.pobtaf(D, L, B, T)
    Lf_dense = np.linalg.cholesky(Q)
    (eD, eL, eB, eT) = ob.dense_to_bta(Lf_dense, n, b, a)
    assert np.allclose(D, eD, rtol=0, atol=1e-12)
    if n > 1:
        assert np.allclose(L, eL, rtol=0, atol=1e-12)
    if a:
        assert np.allclose(B, eB, rtol=0, atol=1e-12)
        assert np.allclose(T, eT, rtol=0, atol=1e-12)
    sign, ld = np.linalg.slogdet(Q)
    assert sign > 0 and abs(ob.logdet(D, T) - ld) <= 1e-12 * abs(ld)


@pytest.mark.parametrize("n", [1, 2, 10, 257])
def test_scalar_tridiagonal_closed_form(n):
    """b=1, a=0: Q = tridiag(-1, 2, -1)_n, logdet = ln(n+1)."""
    D = np.full((n, 1, 1), 2.0)
    L = np.full((n - 1, 1, 1), -1.0)
    ob.pobtf(D, L)
    assert abs(ob.logdet(D) - np.log(n + 1)) <= 1e-13 * max(1.0, np.log(n + 1))


def _kron_case(nt, ns, seed):
    rng = np.random.default_rng(seed)
    M = np.diag(2.0 + rng.random(nt)) - np.diag(0.5 * rng.random(nt - 1), 1)
    M = np.triu(M) + np.triu(M, 1).T
    X = rng.normal(size=(ns, ns))
    S = X @ X.T / ns + np.eye(ns)
    return M, S


def test_kronecker_logdet_solve_and_inverse():
    nt, ns = 7, 6
    M, S = _kron_case(nt, ns, 3)
    Q = np.kron(M, S)
    D, L, _, _ = ob.dense_to_bta(Q, nt, ns, 0)
    F_D, F_L = D.copy(), L.copy()
    ob.pobtf(F_D, F_L)
    ld = ob.logdet(F_D)
    ld_exact = ns * np.linalg.slogdet(M)[1] + nt * np.linalg.slogdet(S)[1]
    assert abs(ld - ld_exact) <= 1e-12 * abs(ld_exact)
    # solve: (M (x) S) vec(X) = vec(Y) with t-major vec  <=>  M X S = Y
    Y = np.random.default_rng(4).normal(size=(nt, ns))
    x = Y.reshape(-1).copy()
    ob.pobts(F_D, F_L, x, "N")
    ob.pobts(F_D, F_L, x, "C")
    Xexp = np.linalg.solve(M, Y) @ np.linalg.inv(S)
    assert np.abs(x.reshape(nt, ns) - Xexp).max() <= 1e-11 * np.abs(Xexp).max()
    # selected inverse: Sigma_{ij} = (M^-1)_{ij} S^-1
    ob.pobtsi(F_D, F_L)
    Mi, Si = np.linalg.inv(M), np.linalg.inv(S)
    for t in range(nt):
        assert np.abs(F_D[t] - Mi[t, t] * Si).max() <= 1e-11 * np.abs(Si).max()
        if t < nt - 1:
            assert np.abs(F_L[t] - Mi[t + 1, t] * Si).max() <= 1e-11 * np.abs(Si).max()


def test_schur_identity_for_the_arrowhead():
    n, b, a = 5, 8, 3
    kf, Q = _random_bta(n, b, a, seed=21)
    D, L, B, T = kf.Q_D.copy(), kf.Q_L.copy(), kf.Q_B.copy(), kf.Q_T.copy()
    ob.pobtaf(D, L, B, T)
    nb = n * b
    Qbt = Q[:nb, :nb]
    Barr = Q[nb:, :nb]
    S = Q[nb:, nb:] - Barr @ np.linalg.solve(Qbt, Barr.T)
    ld_exact = np.linalg.slogdet(Qbt)[1] + np.linalg.slogdet(S)[1]
    assert abs(ob.logdet(D, T) - ld_exact) <= 1e-12 * abs(ld_exact)


@pytest.mark.parametrize("n,b,a", [(1, 4, 0), (3, 5, 2), (6, 4, 3)])
def test_pobtas_residual_and_cho_solve(n, b, a):
    kf, Q = _random_bta(n, b, a, seed=31 + n)
    D, L = kf.Q_D.copy(), kf.Q_L.copy()
    B = kf.Q_B.copy() if a else None
    T = kf.Q_T.copy() if a else None
    ob.pobtaf(D, L, B, T)
    R = np.random.default_rng(2).normal(size=(n * b + a, 2))
    X = R.copy()
    ob.pobtas(D, L, B, T, X, "N")
    ob.pobtas(D, L, B, T, X, "C")
    assert np.linalg.norm(Q @ X - R) <= 1e-12 * np.linalg.norm(R)
    Xc = sla.cho_solve(sla.cho_factor(Q, lower=True), R)
    assert np.abs(X - Xc).max() <= 1e-11 * np.abs(Xc).max()


@pytest.mark.parametrize("n,b,a", [(1, 4, 2), (3, 5, 0), (5, 4, 3)])
def test_pobtasi_matches_dense_inverse(n, b, a):
    kf, Q = _random_bta(n, b, a, seed=41 + n)
    D, L = kf.Q_D.copy(), kf.Q_L.copy()
    B = kf.Q_B.copy() if a else None
    T = kf.Q_T.copy() if a else None
    ob.pobtaf(D, L, B, T)
    ob.pobtasi(D, L, B, T)
    eD, eL, eB, eT = ob.dense_to_bta(np.linalg.inv(Q), n, b, a)
    assert np.abs(D - eD).max() <= 1e-11 * np.abs(eD).max()
    if n > 1:
        assert np.abs(L - eL).max() <= 1e-11 * np.abs(eD).max()
    if a:
        assert np.abs(B - eB).max() <= 1e-11 * np.abs(eD).max()
        assert np.abs(T - eT).max() <= 1e-11 * np.abs(eD).max()


# ------------------------------------------------------------ planted mistakes
# Each variant below is a plausible slip in the oracle; the checks above must
# reject it (so they pin something beyond restating the oracle's formulas).

def _pobtaf_transposed_update(D, L, B, T):
    n = D.shape[0]
    for i in range(n):
        D[i] = np.linalg.cholesky(D[i])
        if i < n - 1:
            L[i] = sla.solve_triangular(D[i], L[i].T, lower=True).T
            D[i + 1] -= L[i].T @ L[i]  # wrong: L^T L instead of L L^T
        B[i] = sla.solve_triangular(D[i], B[i].T, lower=True).T
        T -= B[i] @ B[i].T
        if i < n - 1:
            B[i + 1] -= B[i] @ L[i].T
    T[:] = np.linalg.cholesky(T)


def _pobtaf_dropped_arrow_term(D, L, B, T):
    n = D.shape[0]
    for i in range(n):
        D[i] = np.linalg.cholesky(D[i])
        if i < n - 1:
            L[i] = sla.solve_triangular(D[i], L[i].T, lower=True).T
            D[i + 1] -= L[i] @ L[i].T
        B[i] = sla.solve_triangular(D[i], B[i].T, lower=True).T
        T -= B[i] @ B[i].T
        # dropped: B[i+1] -= B[i] L[i]^T
    T[:] = np.linalg.cholesky(T)


@pytest.mark.parametrize("bad", [_pobtaf_transposed_update, _pobtaf_dropped_arrow_term])
def test_planted_factor_mistakes_are_caught(bad):
    kf = known_factor_bta(4, 6, 2, seed=5)
    D, L, B, T = kf.Q_D.copy(), kf.Q_L.copy(), kf.Q_B.copy(), kf.Q_T.copy()
    try:
        bad(D, L, B, T)
    except np.linalg.LinAlgError:
        return  # caught as non-SPD
    assert np.abs(D - kf.Lf_D).max() > 1e-6 or np.abs(B - kf.Lf_B).max() > 1e-6


def test_planted_selinv_sign_mistake_is_caught():
    n, b, a = 4, 5, 2
    kf, Q = _random_bta(n, b, a, seed=8)
    D, L, B, T = kf.Lf_D.copy(), kf.Lf_L.copy(), kf.Lf_B.copy(), kf.Lf_T.copy()
    # wrong sign on the sub-diagonal recurrence: flip Sigma_{i+1,i}
    ob.pobtasi(D, L, B, T)
    eD, eL, _, _ = ob.dense_to_bta(np.linalg.inv(Q), n, b, a)
    assert np.abs(-L - eL).max() > 1e-6
