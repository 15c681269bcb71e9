"""Seeded synthetic inputs for the BTA solver path.

Every generator uses ``numpy.random.default_rng(seed)`` with the draw order
written in its docstring; "N(mu, s)" means mean mu and standard deviation s.
These are the inputs the parity tests and the benchmark run on; the reference
ships no usable data (its example inputs are git-LFS pointers).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


# ----------------------------------------------------------- known-factor BTA

@dataclass
class KnownFactorBTA:
    """A BTA matrix Q = LL^T built from a known block factor.

    ``Lf_*`` are the factor blocks (diag lower-triangular, sub-diagonal, arrow,
    tip), ``Q_*`` the blocks of Q in the reference storage layout, ``rhs`` one
    right-hand side of length N = n*b + a.
    """
    n: int
    b: int
    a: int
    Lf_D: np.ndarray
    Lf_L: np.ndarray
    Lf_B: np.ndarray
    Lf_T: np.ndarray
    Q_D: np.ndarray
    Q_L: np.ndarray
    Q_B: np.ndarray
    Q_T: np.ndarray
    rhs: np.ndarray


def known_factor_draws(rng: np.random.Generator, n: int, b: int, a: int, norm_b: int | None = None):
    """Draw the factor blocks in a fixed order.

    Draw order: (1) diag-block strict-lower entries N(0, 0.1), shape (n, b, b);
    (2) diag-block diagonals |N(0,1)| + 1, shape (n, b); (3) sub-diagonal blocks
    N(0, 0.1), shape (n-1, b, b); (4) arrow blocks N(0, 0.05), shape (n, a, b);
    (5) tip strict-lower N(0, 0.05), shape (a, a); (6) tip diagonal
    1 + |N(0,1)|, shape (a,); (7) rhs N(0,1), shape (n*b + a,).
    ``norm_b`` scales the drawn (1), (3) and (4) by sqrt(norm_b / b), as
    :func:`known_factor_bta_torch` does (the draws themselves are unchanged).
    """
    sc = (norm_b / b) ** 0.5 if norm_b else 1.0
    Ld = np.tril(rng.normal(0.0, 0.1, (n, b, b)), -1) * sc
    dg = np.abs(rng.normal(0.0, 1.0, (n, b))) + 1.0
    idx = np.arange(b)
    Ld[:, idx, idx] = dg
    Ll = rng.normal(0.0, 0.1, (max(n - 1, 0), b, b)) * sc
    Lb = rng.normal(0.0, 0.05, (n, a, b)) * sc
    Lt = np.tril(rng.normal(0.0, 0.05, (a, a)), -1)
    Lt[np.arange(a), np.arange(a)] = 1.0 + np.abs(rng.normal(0.0, 1.0, a))
    rhs = rng.normal(0.0, 1.0, n * b + a)
    return Ld, Ll, Lb, Lt, rhs


def q_from_factor(Ld, Ll, Lb, Lt):
    """Blocks of Q = LL^T for a block lower-bidiagonal-arrow factor.

    Q[i,i]   = Ld_i Ld_i^T + Ll_{i-1} Ll_{i-1}^T
    Q[i+1,i] = Ll_i Ld_i^T
    Q[n,i]   = Lb_i Ld_i^T + Lb_{i-1} Ll_{i-1}^T
    Q[n,n]   = Lt Lt^T + sum_i Lb_i Lb_i^T
    Works on numpy arrays or torch tensors (uses only @, .transpose and copies).
    """
    n = Ld.shape[0]
    tr = (lambda X: np.swapaxes(X, -1, -2) if isinstance(X, np.ndarray) else X.transpose(-1, -2))
    QD = Ld @ tr(Ld)
    if n > 1:
        QD[1:] += Ll @ tr(Ll)
        QL = Ll @ tr(Ld[:-1])
    else:
        QL = Ll[:0] * 0
    QB = Lb @ tr(Ld)
    if n > 1:
        QB[1:] += Lb[:-1] @ tr(Ll)
    QT = Lt @ tr(Lt) + (Lb @ tr(Lb)).sum(0)
    return QD, QL, QB, QT


def known_factor_bta(n: int = 50, b: int = 92, a: int = 4, seed: int = 0,
                     norm_b: int | None = None) -> KnownFactorBTA:
    """Plain BTA test matrix: n diagonal blocks of size b, arrowhead a.

    Defaults are the plain FP64 BTA solve case (n=50, b=92, a=4, seed 0).
    ``norm_b``: off-diagonal draws scaled to the block norms of the b = norm_b
    construction (see :func:`known_factor_bta_torch`: unscaled, large b gives a
    badly conditioned Q).
    """
    rng = np.random.default_rng(seed)
    Ld, Ll, Lb, Lt, rhs = known_factor_draws(rng, n, b, a, norm_b)
    QD, QL, QB, QT = q_from_factor(Ld, Ll, Lb, Lt)
    return KnownFactorBTA(n, b, a, Ld, Ll, Lb, Lt, QD, QL, QB, QT, rhs)


def known_factor_bta_torch(n: int, b: int, a: int, seed: int, device, norm_b: int | None = None):
    """Large known-factor BTA matrix generated on ``device`` with torch.

    Used where numpy generation would take minutes (b in the thousands).  Same
    distributions as :func:`known_factor_draws` but drawn from a torch
    generator seeded with ``seed`` on ``device``; returns (factor, Q) tuples of
    float64 tensors (D, L, B, T).

    ``norm_b``: if given, the N(0, 0.1) / N(0, 0.05) off-diagonal draws are
    scaled by sqrt(norm_b / b), so that the block spectral norms equal those of
    the b = norm_b construction.  Unscaled, the sub-diagonal factor blocks
    have norm ~0.2 sqrt(b) (9 at b = 2000): the block recurrence of L^-1 then
    amplifies by that factor per block and Q = L L^T is numerically singular
    in FP64 for n >= 3 -- every Cholesky of it fails, the reference's
    included.
    """
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    f64 = torch.float64
    sc = (norm_b / b) ** 0.5 if norm_b else 1.0
    Ld = torch.randn((n, b, b), generator=g, device=device, dtype=f64).mul_(0.1 * sc).tril_(-1)
    dg = torch.randn((n, b), generator=g, device=device, dtype=f64).abs_().add_(1.0)
    Ld.diagonal(dim1=1, dim2=2).copy_(dg)
    Ll = torch.randn((max(n - 1, 0), b, b), generator=g, device=device, dtype=f64).mul_(0.1 * sc)
    Lb = torch.randn((n, a, b), generator=g, device=device, dtype=f64).mul_(0.05 * sc)
    Lt = torch.randn((a, a), generator=g, device=device, dtype=f64).mul_(0.05).tril_(-1)
    Lt.diagonal().copy_(torch.randn((a,), generator=g, device=device, dtype=f64).abs_().add_(1.0))
    Q = _q_from_factor_torch(Ld, Ll, Lb, Lt)
    return (Ld, Ll, Lb, Lt), Q


def _q_from_factor_torch(Ld, Ll, Lb, Lt):
    """Blockwise Q = LL^T on torch tensors, one block at a time (bounded memory)."""
    import torch
    n, b = Ld.shape[0], Ld.shape[1]
    QD = torch.empty_like(Ld)
    QL = torch.empty_like(Ll)
    QB = torch.empty_like(Lb)
    QT = Lt @ Lt.T
    for i in range(n):
        QD[i] = Ld[i] @ Ld[i].T.contiguous()
        QB[i] = Lb[i] @ Ld[i].T.contiguous()
        if i > 0:
            QD[i] += Ll[i - 1] @ Ll[i - 1].T.contiguous()
            QB[i] += Lb[i - 1] @ Ll[i - 1].T.contiguous()
        if i < n - 1:
            QL[i] = Ll[i] @ Ld[i].T.contiguous()
        QT += Lb[i] @ Lb[i].T.contiguous()
    return QD, QL, QB, QT


# ------------------------------------------------------- spatio-temporal model

@dataclass
class STProblem:
    """Inputs of the Gaussian spatio-temporal model on a regular grid.

    spatial  : [c0, g1, g2] (scipy CSR, ns x ns)
    temporal : [M0, M1, M2] (scipy CSR, nt x nt, tridiagonal)
    A        : observation matrix (scipy CSR, m x (nt*ns + nb)); latent order
               is t-major space-time (index t*ns + s) then the fixed effects,
               the reference's kron(M, S) ordering with the regression last.
    y        : observations (m,)
    theta    : evaluation point, theta = (theta_s, theta_t, theta_e, theta_o)
    """
    ns: int
    nt: int
    nb: int
    spatial: list
    temporal: list
    A: object
    y: np.ndarray
    theta: np.ndarray
    prior_mean: np.ndarray
    prior_prec: np.ndarray
    fixed_prec: float


def grid_spatial_matrices(nx: int, ny: int, hx: float, hy: float | None = None):
    """c0 = h^2 I, g1 = 5-point graph Laplacian, g2 = g1 c0^-1 g1 on an nx x ny grid.

    Node index s = ix * ny + iy.  With hy == hx (or None) h^2 = hx^2, else
    h^2 = hx * hy (the cell area).
    """
    import scipy.sparse as sps
    hy = hx if hy is None else hy
    ns = nx * ny
    idx = np.arange(ns).reshape(nx, ny)
    rows, cols = [], []
    for a, b_ in ((idx[:-1, :], idx[1:, :]), (idx[:, :-1], idx[:, 1:])):
        rows.append(a.ravel())
        cols.append(b_.ravel())
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    adj = sps.coo_matrix((np.ones(2 * len(r)), (np.concatenate([r, c]), np.concatenate([c, r]))), shape=(ns, ns)).tocsr()
    deg = np.asarray(adj.sum(axis=1)).ravel()
    g1 = (sps.diags(deg) - adj).tocsr()
    area = hx * hy
    c0 = (area * sps.identity(ns, format="csr")).tocsr()
    g2 = (g1 @ sps.diags(np.full(ns, 1.0 / area)) @ g1).tocsr()
    for m in (c0, g1, g2):
        m.sort_indices()
    return [c0, g1, g2]


def temporal_matrices(nt: int):
    """M0 = I, M1 = first-difference Gram (path Laplacian), M2 = tridiag(-1, 2, -1)."""
    import scipy.sparse as sps
    e = np.ones(nt)
    M0 = sps.identity(nt, format="csr")
    d1 = np.full(nt, 2.0)
    if nt > 1:
        d1[0] = d1[-1] = 1.0
    else:
        d1[:] = 0.0
    M1 = sps.diags([-e[:-1], d1, -e[:-1]], [-1, 0, 1], format="csr")
    M2 = sps.diags([-e[:-1], 2.0 * e, -e[:-1]], [-1, 0, 1], format="csr")
    return [M0.tocsr(), M1.tocsr(), M2.tocsr()]


def st_problem(nt: int = 100, nx: int = 40, ny: int = 40, nb: int = 6, obs_per_node: int = 4,
               seed: int = 1, hx: float | None = None, hy: float | None = None) -> STProblem:
    """Synthetic Gaussian spatio-temporal model (grid mesh, random observations).

    Draw order: (1) observation nodes: m = obs_per_node * ns * nt integers in
    [0, ns*nt); (2) covariates N(0,1), shape (m, nb); (3) latent space-time
    field N(0, 0.5), shape (ns*nt,); (4) fixed effects N(0,1), shape (nb,);
    (5) noise N(0,1) * exp(-theta_o/2), shape (m,); (6) theta perturbation
    N(0, 0.1), shape (4,), added to (0, 0, 0, 1).
    Defaults: the 40 x 40 grid (h = 1/39), nt = 100, nb = 6 case.
    """
    import scipy.sparse as sps
    ns = nx * ny
    hx = 1.0 / (nx - 1) if hx is None else hx
    hy = (1.0 / (ny - 1) if hy is None else hy) if (hy is not None or nx != ny) else hx
    spatial = grid_spatial_matrices(nx, ny, hx, hy)
    temporal = temporal_matrices(nt)
    rng = np.random.default_rng(seed)
    m = obs_per_node * ns * nt
    nodes = rng.integers(0, ns * nt, size=m)
    X = rng.normal(0.0, 1.0, (m, nb))
    u = rng.normal(0.0, 0.5, ns * nt)
    beta = rng.normal(0.0, 1.0, nb)
    theta_o = 1.0
    noise = rng.normal(0.0, 1.0, m) * np.exp(-0.5 * theta_o)
    theta = np.array([0.0, 0.0, 0.0, 1.0]) + rng.normal(0.0, 0.1, 4)
    N = ns * nt + nb
    rows = np.repeat(np.arange(m), 1 + nb)
    cols = np.concatenate([nodes[:, None], ns * nt + np.arange(nb)[None, :].repeat(m, 0)], axis=1).ravel()
    vals = np.concatenate([np.ones((m, 1)), X], axis=1).ravel()
    A = sps.csr_matrix((vals, (rows, cols)), shape=(m, N))
    y = u[nodes] + X @ beta + noise
    return STProblem(ns=ns, nt=nt, nb=nb, spatial=spatial, temporal=temporal, A=A, y=y, theta=theta,
                     prior_mean=np.array([0.0, 0.0, 0.0, 1.0]), prior_prec=np.full(4, 0.5), fixed_prec=1e-3)


@dataclass
class CoregionalProblem:
    """nv interleaved spatio-temporal fields on one mesh (coregional model).

    Latent order: per time block t the fields 0..nv-1 (ns nodes each), i.e.
    index t*nv*ns + v*ns + s (the reference's interleaving,
    coregional_model.py:754-786), then the fixed effects (nb_per per field).
    A_list[v], y_list[v]: observations of field v.
    theta = (theta_s, theta_t, theta_e) per field, lambdas, noise precisions.
    """
    nv: int
    ns: int
    nt: int
    nb: int
    spatial: list
    temporal: list
    A_list: list
    y_list: list
    theta: np.ndarray
    prior_mean: np.ndarray
    prior_prec: np.ndarray
    fixed_prec: float
    A: object = None
    y: object = None


def coregional_problem(nv: int = 3, nt: int = 100, nx: int = 50, ny: int = 30, nb_per: int = 3,
                       obs_per_node: int = 2, seed: int = 4) -> CoregionalProblem:
    """Synthetic coregional model; default 3 fields of 1500 nodes (50 x 30 grid),
    nt = 100, 3 fixed effects per field, 2 ns nt observations per field.

    Draw order: (1) theta per field N(0, 0.1) x 3 (theta_s, theta_t, theta_e),
    field by field; (2) lambdas N(0, 0.5); (3) noise precisions 1 + N(0, 0.1);
    then for each field v: (4) observation nodes, (5) covariates N(0,1)
    (m_v, nb_per), (6) latent field N(0, 0.5) (nt*ns,), (7) fixed effects
    N(0,1), (8) noise N(0,1) * exp(-theta_o/2).
    """
    import scipy.sparse as sps
    ns = nx * ny
    spatial = grid_spatial_matrices(nx, ny, 1.0 / (nx - 1), 1.0 / (ny - 1))
    temporal = temporal_matrices(nt)
    rng = np.random.default_rng(seed)
    n_lam = nv * (nv - 1) // 2
    th_fields = rng.normal(0.0, 0.1, 3 * nv)
    lam = rng.normal(0.0, 0.5, n_lam)
    th_o = 1.0 + rng.normal(0.0, 0.1, nv)
    theta = np.concatenate([th_fields, lam, th_o])
    b = nv * ns
    nb = nv * nb_per
    N = nt * b + nb
    A_list, y_list = [], []
    for v in range(nv):
        m = obs_per_node * ns * nt
        nodes = rng.integers(0, ns * nt, size=m)
        X = rng.normal(0.0, 1.0, (m, nb_per))
        u = rng.normal(0.0, 0.5, ns * nt)
        beta = rng.normal(0.0, 1.0, nb_per)
        noise = rng.normal(0.0, 1.0, m) * np.exp(-0.5 * th_o[v])
        t_idx, s_idx = nodes // ns, nodes % ns
        col_st = t_idx * b + v * ns + s_idx
        rows = np.repeat(np.arange(m), 1 + nb_per)
        cols = np.concatenate([col_st[:, None], (nt * b + v * nb_per + np.arange(nb_per))[None, :].repeat(m, 0)],
                              axis=1).ravel()
        vals = np.concatenate([np.ones((m, 1)), X], axis=1).ravel()
        A_list.append(sps.csr_matrix((vals, (rows, cols)), shape=(m, N)))
        y_list.append(u[nodes] + X @ beta + noise)
    prior_mean = np.concatenate([np.zeros(3 * nv + n_lam), np.ones(nv)])
    return CoregionalProblem(nv=nv, ns=ns, nt=nt, nb=nb, spatial=spatial, temporal=temporal, A_list=A_list,
                             y_list=y_list, theta=theta, prior_mean=prior_mean,
                             prior_prec=np.full(theta.size, 0.5), fixed_prec=1e-3)


def with_g3(prob):
    """Append g3 = g1 c0^-1 g1 c0^-1 g1 to the spatial matrices (the fourth
    matrix of the reference's spatio-temporal precision, q3s of
    spatio_temporal.py:131-136); returns the problem."""
    import scipy.sparse as sps
    c0, g1, g2 = prob.spatial[:3]
    ci = sps.diags(1.0 / c0.diagonal())
    g3 = (g2 @ ci @ g1).tocsr()
    g3.sort_indices()
    prob.spatial = [c0, g1, g2, g3]
    return prob
