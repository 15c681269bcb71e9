"""CPU restatement of f(theta) and its finite-difference gradient (oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows the reference's Gaussian path:
  Q_st      sum of sparse Kronecker products (spatio_temporal.py:114-149,
            here with an arbitrary coefficient table)
  Q_prior   block-diag(Q_st, prec * I_nb) (model.py:306-382, regression.py:28-34)
  Q_cond    Q_prior - A^T (-e^{theta_o} I) A (model.py:384-423, gaussian.py:98-131)
  rhs       -Q_prior x + A^T grad_lik at x = 0, eta = 0 -> e^{theta_o} A^T y
            (model.py:425-449, gaussian.py:64-96)
  f         pyinla.py:602-698 with _evaluate_conditional_latent_parameters
            (:1112-1167) and _evaluate_prior_latent_parameters (:1076-1110)
  FD grad   pyinla.py:512-600
The factorizations use the BTA oracle on blocks scattered as
structured_solver.py:214-250 does.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sps

from . import bta as ob


def q_st(spatial, temporal, table):
    """sum_{m,k} table[m][k] kron(M_m, S_k) as scipy CSR; with a 4-D table
    (nv, nv, nM, nK) the fields are interleaved per time block:
    sum_{v,w,m,k} table[v][w][m][k] kron(M_m, kron(E_vw, S_k))."""
    table = np.asarray(table)
    if table.ndim == 2:
        table = table[None, None]
    nv = table.shape[0]
    Q = None
    for v in range(nv):
        for w in range(nv):
            E = sps.csr_matrix(([1.0], ([v], [w])), shape=(nv, nv))
            for m, M in enumerate(temporal):
                for k, S in enumerate(spatial):
                    if k >= table.shape[3] or table[v, w, m, k] == 0.0:
                        continue
                    term = table[v, w, m, k] * sps.kron(M, sps.kron(E, S), format="csr")
                    Q = term if Q is None else Q + term
    return Q.tocsr()


def _obs(prob):
    """[(A_v, y_v, theta index of its noise precision)] (one set for nv = 1).
    The indices are prob.noise_idx when the problem gives them, else the
    last components of theta (one per observation set)."""
    if getattr(prob, "A_list", None) is not None:
        idx = getattr(prob, "noise_idx", None)
        if idx is None:
            d_q = len(prob.theta) - len(prob.A_list)
            idx = [d_q + v for v in range(len(prob.A_list))]
        return [(sps.csr_matrix(A), y, idx[v]) for v, (A, y) in enumerate(zip(prob.A_list, prob.y_list))]
    idx = getattr(prob, "noise_idx", None)
    return [(sps.csr_matrix(prob.A), prob.y, 3 if idx is None else idx[0])]


def q_prior(prob, table):
    nK = np.asarray(table).shape[-1]
    Qst = q_st(prob.spatial[:nK], prob.temporal, table)
    return sps.block_diag([Qst, prob.fixed_prec * sps.identity(prob.nb)], format="csr")


def q_cond(prob, Qp, theta):
    Q = Qp
    for A, y, idx in _obs(prob):
        Q = Q + np.exp(theta[idx]) * (A.T @ A)
    return Q.tocsr()


def scatter_bta(Q, n, b, a, sparsity="bta"):
    """Dense BTA blocks of sparse Q (structured_solver.py:214-250)."""
    Qc = sps.csc_matrix(Q)
    D = np.zeros((n, b, b))
    L = np.zeros((max(n - 1, 0), b, b))
    B = np.zeros((n, a, b)) if sparsity == "bta" else None
    T = np.zeros((a, a)) if sparsity == "bta" else None
    for i in range(n):
        D[i] = Qc[i * b:(i + 1) * b, i * b:(i + 1) * b].toarray()
        if i < n - 1:
            L[i] = Qc[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b].toarray()
        if sparsity == "bta":
            B[i] = Qc[n * b:, i * b:(i + 1) * b].toarray()
    if sparsity == "bta":
        T[:] = Qc[n * b:, n * b:].toarray()
    return D, L, B, T


def log_prior(theta, mean, prec):
    return float(np.sum(-0.5 * prec * (np.asarray(theta) - mean) ** 2))


def _log_prior(prob, theta):
    """log p(theta): the problem's prior objects when it has them (the PC and
    Gaussian priors, pinned by tests/test_priors.py), else independent
    Gaussians (prior_hyperparameters/gaussian.py:26-29)."""
    pr = getattr(prob, "priors", None)
    if pr is not None:
        return float(pr.log_prior(theta))
    return log_prior(theta, prob.prior_mean, prob.prior_prec)


def evaluate_f(prob, table_fn, theta, return_parts=False):
    theta = np.asarray(theta, dtype=np.float64)
    n, b, a = prob.nt, getattr(prob, "nv", 1) * prob.ns, prob.nb
    table = table_fn(theta)
    Qp = q_prior(prob, table)
    Qc = q_cond(prob, Qp, theta)
    obs = _obs(prob)
    rhs = sum(np.exp(theta[idx]) * (A.T @ y) for A, y, idx in obs)
    D, L, B, T = scatter_bta(Qc, n, b, a, "bta")
    ob.pobtaf(D, L, B, T)
    x = rhs.copy()
    ob.pobtas(D, L, B, T, x, "N")
    ob.pobtas(D, L, B, T, x, "C")
    ld_c = ob.logdet(D, T)
    quad = float(x @ (Qc @ x))
    Dp, Lp, _, _ = scatter_bta(Qp, n, b, a, "bt")
    ob.pobtf(Dp, Lp)
    ld_p = ob.logdet(Dp)
    logp = _log_prior(prob, theta)
    lik = sum(0.5 * theta[idx] * A.shape[0] - 0.5 * np.exp(theta[idx]) * float(y @ y) for A, y, idx in obs)
    f = (0.5 * ld_c - 0.5 * quad) - (logp + lik + 0.5 * ld_p)
    if return_parts:
        return dict(f=f, ld_c=ld_c, ld_p=ld_p, quad=quad, logp=logp, lik=lik, x=x)
    return f


def marginal_variances(prob, table_fn, theta):
    """Marginal variances of the latent parameters and of the observations
    (PyINLA.get_marginal_variances_latent_parameters /
    get_marginal_variances_observations, pyinla.py:912-1016, Gaussian branch):
    Q_cond(theta) factored and inverted on its BTA pattern (pobtaf, pobtasi);
    diag(Sigma), then diag(A Sigma_pattern A^T) with Sigma kept on the pattern
    of Q_cond (as _structured_to_spmatrix(Q_conditional) does,
    structured_solver.py:470-538).  Returns (latent, [per observation set])."""
    theta = np.asarray(theta, dtype=np.float64)
    n, b, a = prob.nt, getattr(prob, "nv", 1) * prob.ns, prob.nb
    Qc = q_cond(prob, q_prior(prob, table_fn(theta)), theta)
    D, L, B, T = scatter_bta(Qc, n, b, a, "bta")
    ob.pobtaf(D, L, B, T)
    ob.pobtasi(D, L, B, T)
    latent = np.concatenate([np.concatenate([np.diag(D[i]) for i in range(n)]), np.diag(T)])
    S = sps.csr_matrix(ob.bta_to_dense(D, L, B, T))
    pat = sps.csr_matrix(Qc)
    pat.data[:] = 1.0
    S = S.multiply(pat).tocsr()  # Sigma on the pattern of Q_cond
    obs = [np.asarray((A @ S @ A.T).diagonal()).ravel() for A, _, _ in _obs(prob)]
    return latent, obs


def hessian_fd(prob, table_fn, theta, eps=5e-3):
    """Central-difference Hessian of f with the stencil of
    PyINLA._evaluate_hessian_f (pyinla.py:767-910)."""
    theta = np.asarray(theta, dtype=np.float64)
    d = theta.size
    E = eps * np.eye(d)
    f0 = evaluate_f(prob, table_fn, theta)
    H = np.zeros((d, d))
    for i in range(d):
        H[i, i] = (evaluate_f(prob, table_fn, theta + E[i]) - 2.0 * f0
                   + evaluate_f(prob, table_fn, theta - E[i])) / eps ** 2
        for j in range(i + 1, d):
            fpp = evaluate_f(prob, table_fn, theta + E[i] + E[j])
            fpm = evaluate_f(prob, table_fn, theta + E[i] - E[j])
            fmp = evaluate_f(prob, table_fn, theta - E[i] + E[j])
            fmm = evaluate_f(prob, table_fn, theta - E[i] - E[j])
            H[i, j] = H[j, i] = (fpp - fpm - fmp + fmm) / (4.0 * eps * eps)
    return H


def fd_gradient(prob, table_fn, theta, eps=1e-3):
    """Central FD with the reference's evaluation order (pyinla.py:540-580)."""
    theta = np.asarray(theta, dtype=np.float64)
    d = theta.size
    mat = np.repeat(theta[:, None], 2 * d + 1, axis=1)
    mat[:, 1:1 + d] += eps * np.eye(d)
    mat[:, d + 1:] -= eps * np.eye(d)
    fv = np.zeros(2 * d + 1)
    for k in range(2 * d, -1, -1):
        fv[k] = evaluate_f(prob, table_fn, mat[:, k])
    return fv[0], (fv[1:1 + d] - fv[d + 1:]) / (2 * eps)


def dense_f_torch(prob, table_fn_torch, theta):
    """f(theta) from dense float64 torch linear algebra (differentiable).

    Independent of the BTA recurrences: dense Cholesky of the full matrices,
    used to pin the adjoint gradient by autograd at small sizes.  table_fn_torch
    returns a (nM, nK) table or an (nv, nv, nM, nK) coregional table.
    """
    import torch
    th = theta
    nv = getattr(prob, "nv", 1)
    n, b, a = prob.nt, nv * prob.ns, prob.nb
    table = table_fn_torch(th)
    if table.dim() == 2:
        table = table[None, None]
    nK = table.shape[3]
    Sd = [torch.as_tensor(s.toarray()) for s in prob.spatial[:nK]]
    Md = [torch.as_tensor(sps.csr_matrix(M).toarray()) for M in prob.temporal]
    N = n * b + a
    Qst = torch.zeros((n * b, n * b), dtype=torch.float64)
    for v in range(nv):
        for w in range(nv):
            E = torch.zeros((nv, nv), dtype=torch.float64)
            E[v, w] = 1.0
            for m, M in enumerate(Md):
                for k, S in enumerate(Sd):
                    Qst = Qst + table[v, w, m, k] * torch.kron(M, torch.kron(E, S))
    Qp = torch.zeros((N, N), dtype=torch.float64)
    Qp[: n * b, : n * b] = Qst
    Qp[n * b:, n * b:] = prob.fixed_prec * torch.eye(a, dtype=torch.float64)
    Qc = Qp
    rhs = torch.zeros(N, dtype=torch.float64)
    lik = torch.zeros((), dtype=torch.float64)
    for A_s, y_n, idx in _obs(prob):
        A = torch.as_tensor(A_s.toarray())
        y = torch.as_tensor(np.asarray(y_n))
        eo = torch.exp(th[idx])
        Qc = Qc + eo * (A.T @ A)
        rhs = rhs + eo * (A.T @ y)
        lik = lik + 0.5 * th[idx] * A.shape[0] - 0.5 * eo * (y @ y)
    Lc = torch.linalg.cholesky(Qc)
    x = torch.cholesky_solve(rhs[:, None], Lc)[:, 0]
    ld_c = 2 * torch.log(torch.diagonal(Lc)).sum()
    Lp = torch.linalg.cholesky(Qst)
    ld_p = 2 * torch.log(torch.diagonal(Lp)).sum()
    quad = x @ (Qc @ x)
    logp = _log_prior_torch(prob, th)
    return (0.5 * ld_c - 0.5 * quad) - (logp + lik + 0.5 * ld_p)


def _log_prior_torch(prob, th):
    """log p(theta) in torch (differentiable): Gaussian, or the PC priors of
    penalized_complexity.py:47-85 by hyperparameter type."""
    import torch
    pr = getattr(prob, "priors", None)
    if pr is None:
        mean = torch.as_tensor(prob.prior_mean)
        prec = torch.as_tensor(prob.prior_prec)
        return torch.sum(-0.5 * prec * (th - mean) ** 2)
    total = torch.zeros((), dtype=torch.float64)
    for i, p in enumerate(pr.priors):
        t = th[i]
        kind = getattr(p, "kind", "gaussian")
        if kind == "gaussian":
            total = total - 0.5 * p.precision * (t - p.mean) ** 2
        elif kind == "r_s":
            total = total + np.log(p.lam) - p.lam * torch.exp(-t) - t
        elif kind == "r_t":
            total = total + np.log(p.lam) - p.lam * torch.exp(-0.5 * t) + np.log(0.5) - 0.5 * t
        elif kind in ("sigma_st", "sigma_e", "prec_o"):
            total = total + np.log(p.lam) - p.lam * torch.exp(t) + t
    return total


def st_q_reference(spatial, temporal, r_s, r_t, sigma_st, gammas):
    """One field's spatio-temporal precision as SpatioTemporalSubModel
    builds it (spatio_temporal.py:114-149): with (gamma_s, gamma_t,
    gamma_st) = gammas(r_s, r_t, sigma_st),
      q1s = e^{2gs} c0 + g1, q2s = e^{4gs} c0 + 2 e^{2gs} g1 + g2,
      q3s = e^{6gs} c0 + 3 e^{4gs} g1 + 3 e^{2gs} g2 + g3,
      Q = e^{2gst} (m0 (x) q3s + e^{gt} m1 (x) q2s + e^{2gt} m2 (x) q1s)."""
    c0, g1, g2, g3 = spatial[:4]
    m0, m1, m2 = temporal
    gs, gt, gst = (float(g) for g in gammas(r_s, r_t, sigma_st))
    es, et, est = np.exp(gs), np.exp(gt), np.exp(gst)
    q1s = es ** 2 * c0 + g1
    q2s = es ** 4 * c0 + 2 * es ** 2 * g1 + g2
    q3s = es ** 6 * c0 + 3 * es ** 4 * g1 + 3 * es ** 2 * g2 + g3
    return (est ** 2 * (sps.kron(m0, q3s) + et * sps.kron(m1, q2s) + et ** 2 * sps.kron(m2, q1s))).tocsc()


def coregional_q_prior_reference(prob, theta, sigma_st, gammas):
    """Q_prior of CoregionalModel.construct_Q_prior (coregional_model.py:274-555)
    restated block by block: per-field Qu_v, the sigma / lambda blocks
    q_vw (:296-408), the field-major block matrix permuted to the
    time-interleaved order [t][field][node] (:754-786), then the fixed-effects
    tip.  theta layout: (r_s, r_t, prec_o) per field, sigmas, lambdas."""
    nv, ns, nt = prob.nv, prob.ns, prob.nt
    th = np.asarray(theta, dtype=np.float64)
    Qu = [st_q_reference(prob.spatial, prob.temporal, th[3 * v], th[3 * v + 1], sigma_st[v], gammas)
          for v in range(nv)]
    s2 = [np.exp(th[3 * nv + v]) ** 2 for v in range(nv)]
    lam = th[4 * nv:]
    if nv == 2:
        l01 = lam[0]
        q = {(0, 0): Qu[0] / s2[0] + l01 ** 2 / s2[1] * Qu[1], (1, 0): -l01 / s2[1] * Qu[1],
             (1, 1): Qu[1] / s2[1]}
    else:
        l01, l02, l12 = lam
        q = {(0, 0): Qu[0] / s2[0] + l01 ** 2 / s2[1] * Qu[1] + l12 ** 2 / s2[2] * Qu[2],
             (1, 0): -l01 / s2[1] * Qu[1] + l02 * l12 / s2[2] * Qu[2],
             (2, 0): -l12 / s2[2] * Qu[2],
             (1, 1): Qu[1] / s2[1] + l02 ** 2 / s2[2] * Qu[2],
             (2, 1): -l02 / s2[2] * Qu[2],
             (2, 2): Qu[2] / s2[2]}
    blocks = [[None] * nv for _ in range(nv)]
    for (v, w), m in q.items():
        blocks[v][w] = m
        if v != w:
            blocks[w][v] = m.T
    Qfm = sps.bmat(blocks, format="csr")  # field-major: index v*nt*ns + t*ns + s
    n_re = nt * ns
    perm = np.empty(nv * n_re, dtype=np.int64)  # interleaved position -> field-major index
    for t in range(nt):
        for v in range(nv):
            perm[(t * nv + v) * ns:(t * nv + v + 1) * ns] = v * n_re + t * ns + np.arange(ns)
    Qst = Qfm[perm][:, perm]
    return sps.block_diag([Qst, prob.fixed_prec * sps.identity(prob.nb)], format="csr")
