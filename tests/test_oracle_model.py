"""Pin the f(theta) oracle (oracle/model.py) on the CPU.

f is pinned by the exact Gaussian marginal likelihood (f + log p(theta) +
log N(y; 0, A Q_p^-1 A^T + D_o^-1) is a known theta-free constant, with the
marginal density formed from a dense m x m covariance), its FD gradient by the
autograd gradient of that marginal, and the structure of f by the INLA
identity at several x.  It is also checked against a dense float64 torch
restatement of the same formula.  Coefficient tables are checked against
hand-expanded values.
"""
import numpy as np
import pytest
import scipy.sparse as sps
import torch

from oracle import model as om
from pyinla_amd.coefficients import (CoefficientModel, CoregionalCoefficients, baseline_table,
                                     coregional_weights, reference_table)
from pyinla_amd.workloads import coregional_problem, st_problem


def test_st_f_and_fd_against_dense_autograd():
    p = st_problem(nt=6, nx=5, ny=4, nb=3, seed=2)
    cm = CoefficientModel("baseline")
    f0, g = om.fd_gradient(p, cm.table, p.theta, eps=1e-4)
    th = torch.tensor(p.theta, requires_grad=True)
    fd = om.dense_f_torch(p, baseline_table, th)
    fd.backward()
    assert abs(f0 - float(fd.detach())) <= 1e-12 * abs(f0)
    assert np.abs(th.grad.numpy() - g).max() <= 1e-6 * np.abs(g).max()


def test_coregional_f_and_fd_against_dense_autograd():
    p = coregional_problem(nv=3, nt=3, nx=4, ny=3, nb_per=2, obs_per_node=2, seed=1)
    c = CoregionalCoefficients(3)
    f0, g = om.fd_gradient(p, c.table, p.theta, eps=1e-4)
    th = torch.tensor(p.theta, requires_grad=True)
    fd = om.dense_f_torch(p, c._fn, th)
    fd.backward()
    assert abs(f0 - float(fd.detach())) <= 1e-12 * abs(f0)
    assert np.abs(th.grad.numpy() - g).max() <= 1e-6 * np.abs(g).max()


def _dense_model(p, table):
    """Dense Q_prior (space-time part and fixed-effects tip), the stacked
    observation matrix and the theta index of each observation row's noise
    precision."""
    Qp = om.q_prior(p, table).toarray()
    obs = om._obs(p)
    A = np.vstack([Av.toarray() for Av, _, _ in obs])
    y = np.concatenate([np.asarray(yv) for _, yv, _ in obs])
    idx = np.concatenate([np.full(Av.shape[0], i) for Av, _, i in obs])
    return Qp, A, y, idx


def _log_normal(x, mean, prec=None, cov=None):
    """log N(x; mean, .) from a dense precision or covariance (all 2 pi terms)."""
    d = x - mean
    k = x.size
    if prec is not None:
        sign, ld = np.linalg.slogdet(prec)
        return 0.5 * ld - 0.5 * d @ prec @ d - 0.5 * k * np.log(2 * np.pi)
    sign, ld = np.linalg.slogdet(cov)
    return -0.5 * ld - 0.5 * d @ np.linalg.solve(cov, d) - 0.5 * k * np.log(2 * np.pi)


def _problems():
    st = st_problem(nt=4, nx=4, ny=4, nb=3, seed=3)
    co = coregional_problem(nv=2, nt=3, nx=4, ny=3, nb_per=2, obs_per_node=2, seed=5)
    return [("st", st, CoefficientModel("baseline")), ("coregional", co, CoregionalCoefficients(2))]


@pytest.mark.parametrize("case", [0, 1])
def test_f_is_the_exact_marginal_likelihood(case):
    """For a Gaussian likelihood the INLA objective is exact: with every
    constant kept,
        f(theta) + log p(theta) + log N(y; 0, A Q_p^-1 A^T + D_o^-1)
          = 1/2 nb ln(prec_fixed) - 1/2 m ln(2 pi),
    the right side collecting the 2 pi terms f drops and the fixed-effects
    tip the BT prior log-determinant leaves out (D_o = diag of the noise
    precisions e^{theta_o} per observation row).  The marginal density is
    formed from the m x m dense covariance, independently of any term of f,
    and the identity must hold at several theta."""
    name, p, cm = _problems()[case]
    rng = np.random.default_rng(11)
    const = 0.5 * p.nb * np.log(p.fixed_prec)

    def residual(fval, th):
        Qp, A, y, idx = _dense_model(p, cm.table(th))
        Sy = A @ np.linalg.inv(Qp) @ A.T + np.diag(np.exp(-th[idx]))
        logp = om.log_prior(th, p.prior_mean, p.prior_prec)
        return fval + logp + _log_normal(y, np.zeros_like(y), cov=Sy) - (const - 0.5 * y.size * np.log(2 * np.pi))

    for k in range(3):
        th = np.asarray(p.theta) + (0 if k == 0 else rng.normal(0, 0.2, len(p.theta)))
        f = om.evaluate_f(p, cm.table, th)
        assert abs(residual(f, th)) <= 1e-10 * abs(f), (name, k)
    # planted mistakes in the terms of f break the identity: the prior
    # logdet with the fixed-effects tip, a flipped quadratic term, the
    # likelihood's theta_o term dropped
    th = np.asarray(p.theta)
    pt = om.evaluate_f(p, cm.table, th, return_parts=True)
    m = sum(Av.shape[0] for Av, _, _ in om._obs(p))
    for bad in (pt["f"] - 0.5 * p.nb * np.log(p.fixed_prec), pt["f"] + pt["quad"],
                pt["f"] + 0.5 * th[-1] * m / len(om._obs(p))):
        assert abs(residual(bad, th)) > 1e-3 * abs(pt["f"]) or abs(residual(bad, th)) > 1.0


@pytest.mark.parametrize("case", [0, 1])
def test_fd_gradient_is_the_gradient_of_the_marginal_likelihood(case):
    """The reference-order central FD gradient of the oracle f equals the
    autograd gradient of -log p(theta) - log N(y; 0, A Q_p(theta)^-1 A^T +
    D_o(theta)^-1), built densely in torch from the coefficient table alone."""
    name, p, cm = _problems()[case]
    th0 = np.asarray(p.theta, dtype=np.float64)
    _, g_fd = om.fd_gradient(p, cm.table, th0, eps=1e-4)
    fn = cm.fn if hasattr(cm, "fn") else cm._fn
    nv = getattr(p, "nv", 1)
    obs = om._obs(p)
    A = torch.as_tensor(np.vstack([Av.toarray() for Av, _, _ in obs]))
    y = torch.as_tensor(np.concatenate([np.asarray(yv) for _, yv, _ in obs]))
    idx = torch.as_tensor(np.concatenate([np.full(Av.shape[0], i) for Av, _, i in obs]))
    Sd = [torch.as_tensor(s_.toarray()) for s_ in p.spatial]
    Md = [torch.as_tensor(sps.csr_matrix(M).toarray()) for M in p.temporal]

    def neg_log_marginal(t):
        tab = fn(t[: len(th0) - len(obs) if nv > 1 else 3])
        if tab.dim() == 2:
            tab = tab[None, None]
        n, b = p.nt, nv * p.ns
        Qst = torch.zeros((n * b, n * b), dtype=torch.float64)
        for v in range(nv):
            for w in range(nv):
                E = torch.zeros((nv, nv), dtype=torch.float64)
                E[v, w] = 1.0
                for m_, M in enumerate(Md):
                    for k_, S in enumerate(Sd[: tab.shape[3]]):
                        Qst = Qst + tab[v, w, m_, k_] * torch.kron(M, torch.kron(E, S))
        Qp = torch.block_diag(Qst, p.fixed_prec * torch.eye(p.nb, dtype=torch.float64))
        Sy = A @ torch.linalg.solve(Qp, A.T) + torch.diag(torch.exp(-t[idx]))
        Ly = torch.linalg.cholesky(Sy)
        z = torch.linalg.solve_triangular(Ly, y[:, None], upper=False)[:, 0]
        logN = -torch.log(torch.diagonal(Ly)).sum() - 0.5 * z @ z
        logp = torch.sum(-0.5 * torch.as_tensor(p.prior_prec) * (t - torch.as_tensor(p.prior_mean)) ** 2)
        return -logp - logN

    t = torch.tensor(th0, requires_grad=True)
    neg_log_marginal(t).backward()
    g = t.grad.numpy()
    assert np.abs(g - g_fd).max() <= 1e-6 * np.abs(g).max(), (name, g, g_fd)


def test_inla_identity_holds_at_every_x():
    """log p(theta | y) up to theta-free constants is
        log p(theta) + log p(x | theta) + log p(y | x, theta) - log p(x | theta, y)
    at ANY x (the Gaussian conditional is exact).  The reference evaluates it
    at x = 0 (pyinla.py:640-641); evaluated with full densities at x = 0,
    x = x* and a random x it must give the same value, and that value is
    -f(theta) up to the constant of test_f_is_the_exact_marginal_likelihood."""
    p = st_problem(nt=4, nx=4, ny=4, nb=2, seed=3)
    cm = CoefficientModel()
    th = np.asarray(p.theta)
    parts = om.evaluate_f(p, cm.table, th, return_parts=True)
    Qp, A, y, idx = _dense_model(p, cm.table(th))
    eo = np.exp(th[idx])
    Qc = Qp + A.T @ (eo[:, None] * A)
    xs = np.linalg.solve(Qc, A.T @ (eo * y))
    assert np.abs(xs - parts["x"]).max() <= 1e-10 * np.abs(xs).max()
    logp = om.log_prior(th, p.prior_mean, p.prior_prec)
    vals = []
    for x in (np.zeros_like(xs), xs, np.random.default_rng(2).normal(size=xs.size)):
        lp_x = _log_normal(x, np.zeros_like(x), prec=Qp)
        lp_y = np.sum(0.5 * np.log(eo) - 0.5 * eo * (y - A @ x) ** 2 - 0.5 * np.log(2 * np.pi))
        lp_cond = _log_normal(x, xs, prec=Qc)
        vals.append(logp + lp_x + lp_y - lp_cond)
    assert max(vals) - min(vals) <= 1e-9 * abs(vals[0])
    const = 0.5 * p.nb * np.log(p.fixed_prec) - 0.5 * y.size * np.log(2 * np.pi)
    assert abs(-parts["f"] + const - vals[0]) <= 1e-10 * abs(parts["f"])
    # the quadratic term equals x*^T b at x* = Q_c^-1 b
    rhs = np.exp(th[3]) * (sps.csr_matrix(p.A).T @ p.y)
    assert abs(parts["quad"] - float(xs @ rhs)) <= 1e-10 * abs(parts["quad"])


def test_baseline_table_hand_expanded():
    ts, tt, te = 0.3, -0.2, 0.1
    t = baseline_table(torch.tensor([ts, tt, te], dtype=torch.float64)).numpy()
    e = np.exp
    exp = np.array([[e(te + 2 * tt + 4 * ts), 2 * e(te + 2 * tt + 2 * ts), e(te + 2 * tt)],
                    [2 * e(te + tt + 2 * ts), 2 * e(te + tt), 0.0],
                    [e(te), 0.0, 0.0]])
    np.testing.assert_allclose(t, exp, rtol=1e-15)


def test_reference_table_plane_closed_form():
    """gamma transform of spatio_temporal.py:151-197 (plane): alpha = 2,
    nu_s = 1, nu_t = 1/2, c1 = Gamma(1/2) Gamma(1) / (Gamma(1) Gamma(2) (4 pi)^1.5)."""
    rs, rt, sst = 0.4, -0.1, 0.2
    t = reference_table(torch.tensor([rs, rt, sst], dtype=torch.float64), "plane").numpy()
    gs = 0.5 * np.log(8.0) - rs
    gt = rt - 0.5 * np.log(4.0) + 2 * gs
    c1 = np.sqrt(np.pi) / (4 * np.pi) ** 1.5
    gst = 0.5 * np.log(c1) - 0.5 * gt - gs - sst
    es = np.exp(gs)
    s = np.exp(2 * gst)
    np.testing.assert_allclose(t[0], s * np.array([es**6, 3 * es**4, 3 * es**2, 1.0]), rtol=1e-14)
    np.testing.assert_allclose(t[1], s * np.exp(gt) * np.array([es**4, 2 * es**2, 1.0, 0.0]), rtol=1e-14)
    np.testing.assert_allclose(t[2], s * np.exp(2 * gt) * np.array([es**2, 1.0, 0.0, 0.0]), rtol=1e-14)


@pytest.mark.parametrize("nv", [2, 3])
def test_coregional_weights_are_a_congruence(nv):
    """sum_u C[v][w][u] d_u equals (L^-T diag(d) L^-1)[v][w] for the unit
    lower-triangular L^-1 the weights encode (so Q is SPD for any lambdas)."""
    rng = np.random.default_rng(0)
    lam = rng.normal(size=nv * (nv - 1) // 2)
    d = rng.random(nv) + 0.5
    C = coregional_weights(torch.tensor(lam), nv).numpy()
    Q = np.einsum("vwu,u->vw", C, d)
    Li = np.eye(nv)
    if nv == 2:
        Li[1, 0] = -lam[0]
    else:
        Li[1, 0], Li[2, 0], Li[2, 1] = -lam[0], -lam[2], -lam[1]
    np.testing.assert_allclose(Q, Li.T @ np.diag(d) @ Li, rtol=1e-14)


def test_workload_generators_are_seeded():
    a = st_problem(nt=3, nx=4, ny=4, nb=2, seed=9)
    b = st_problem(nt=3, nx=4, ny=4, nb=2, seed=9)
    assert (a.A != b.A).nnz == 0 and np.array_equal(a.y, b.y) and np.array_equal(a.theta, b.theta)


@pytest.mark.parametrize("kind", ["st", "coregional"])
def test_marginal_variances_against_dense_inverse(kind):
    """diag(Sigma) and diag(A Sigma A^T) from the BTA selected inversion equal
    the same quantities from the dense inverse of Q_cond (every pair of
    columns of an observation row lies on the pattern of A^T A)."""
    if kind == "st":
        p = st_problem(nt=5, nx=4, ny=4, nb=3, seed=4)
        cm = CoefficientModel("baseline")
    else:
        p = coregional_problem(nv=2, nt=3, nx=4, ny=3, nb_per=2, obs_per_node=2, seed=3)
        cm = CoregionalCoefficients(2)
    lat, obs = om.marginal_variances(p, cm.table, p.theta)
    Qc = om.q_cond(p, om.q_prior(p, cm.table(np.asarray(p.theta))), np.asarray(p.theta))
    S = np.linalg.inv(Qc.toarray())
    assert np.abs(lat - np.diag(S)).max() <= 1e-10 * np.abs(np.diag(S)).max()
    for (A, _, _), v in zip(om._obs(p), obs):
        ref = np.einsum("ij,jk,ik->i", A.toarray(), S, A.toarray())
        assert np.abs(v - ref).max() <= 1e-10 * np.abs(ref).max()


def test_hessian_fd_against_dense_autograd():
    """The reference-stencil FD Hessian of the oracle f against the exact
    Hessian of the dense torch f (second-order autograd); the FD truncation
    error at eps = 5e-3 bounds the agreement."""
    p = st_problem(nt=5, nx=4, ny=4, nb=3, seed=6)
    cm = CoefficientModel("baseline")
    H = om.hessian_fd(p, cm.table, p.theta, eps=5e-3)
    th = torch.tensor(p.theta, dtype=torch.float64)
    Hx = torch.autograd.functional.hessian(lambda t: om.dense_f_torch(p, baseline_table, t), th).numpy()
    assert np.allclose(H, H.T)
    # O(eps^2) truncation: 1.4e-6 at eps = 5e-3, a quarter of it at half the step
    err = np.abs(H - Hx).max() / np.abs(Hx).max()
    assert err <= 1e-5
    H2 = om.hessian_fd(p, cm.table, p.theta, eps=2.5e-3)
    assert np.abs(H2 - Hx).max() / np.abs(Hx).max() <= 0.4 * err
