"""Pin the hyperparameter priors, the theta transforms and the reference
coregional parameterization on the CPU.

* Penalized-complexity priors (penalized_complexity.py:28-85) are proper
  densities in theta with P(range < u) = alpha (r_s, r_t) and
  P(scale > u) = alpha (sigma_st, sigma_e, prec_o): checked by quadrature
  and by the closed-form tail probability; gradients against central
  differences.
* model_to_interpretable inverts interpretable_to_model on both manifolds
  (spatio_temporal.py:151-263); the plane transform has a closed form.
* The coregional table of ReferenceCoregionalCoefficients gives the same
  Q_prior as a block-by-block restatement of CoregionalModel.construct_Q_prior
  (oracle/model.py: coregional_q_prior_reference).
* With PC priors in theta, f still satisfies the exact Gaussian-marginal
  identity (the prior enters f with the right sign).
"""
import math

import numpy as np
import pytest
import scipy.integrate as si
import scipy.sparse as sps

from oracle import model as om
from pyinla_amd.coefficients import (ReferenceCoregionalCoefficients, interpretable_to_model,
                                     model_to_interpretable)
from pyinla_amd.priors import GaussianPrior, HyperPriors, PenalizedComplexityPrior
from pyinla_amd.workloads import coregional_problem, with_g3


@pytest.mark.parametrize("kind,alpha,u", [("r_s", 0.01, 0.5), ("r_s", 0.05, 2.0), ("r_t", 0.01, 5.0),
                                          ("r_t", 0.1, 0.3), ("sigma_st", 0.01, 3.0), ("sigma_e", 0.05, 1.0),
                                          ("prec_o", 0.01, 5.0)])
def test_pc_prior_is_a_density_with_the_stated_tail(kind, alpha, u):
    p = PenalizedComplexityPrior(alpha, u, kind)
    dens = lambda t: math.exp(p.log_prior(t))
    total = si.quad(dens, -60, 60, limit=400, points=[math.log(u)])[0]
    assert abs(total - 1.0) < 1e-8
    if kind in ("r_s", "r_t"):
        tail = si.quad(dens, -60, math.log(u), limit=400)[0]  # P(range < u)
    else:
        tail = si.quad(dens, math.log(u), 60, limit=400)[0]   # P(scale > u)
    assert abs(tail - alpha) < 1e-8
    for t in (-1.3, 0.0, 0.7, 2.1):
        h = 1e-5
        fd = (p.log_prior(t + h) - p.log_prior(t - h)) / (2 * h)
        assert abs(p.grad(t) - fd) <= 1e-7 * max(1.0, abs(fd))


def test_untyped_pc_prior_is_zero_as_in_the_reference():
    """CoregionalModel builds the priors of its sigmas without a type
    (coregional_model.py:160-164): lambda = 0 and log p = 0."""
    p = PenalizedComplexityPrior(0.01, 1.0, None)
    assert p.lam == 0.0 and p.log_prior(0.3) == 0.0 and p.grad(0.3) == 0.0


def test_gaussian_prior_and_collection():
    g = GaussianPrior(0.5, 2.0)
    assert g.log_prior(1.5) == pytest.approx(-1.0)
    assert g.grad(1.5) == pytest.approx(-2.0)
    hp = HyperPriors([g, PenalizedComplexityPrior(0.01, 2.0, "r_s")])
    th = np.array([1.5, 0.2])
    assert hp.log_prior(th) == pytest.approx(g.log_prior(1.5) + hp.priors[1].log_prior(0.2))
    np.testing.assert_allclose(hp.grad(th), [g.grad(1.5), hp.priors[1].grad(0.2)])
    with pytest.raises(ValueError):
        hp.log_prior(np.zeros(3))


@pytest.mark.parametrize("manifold", ["plane", "sphere"])
def test_theta_transform_round_trip(manifold):
    rng = np.random.default_rng(3)
    for _ in range(5):
        th = rng.normal(0, 0.8, 3)
        g = interpretable_to_model(*th, manifold=manifold)
        back = model_to_interpretable(*[float(x) for x in g], manifold=manifold)
        np.testing.assert_allclose(back, th, atol=1e-13)


def test_theta_transform_plane_closed_form():
    """model_to_interpretable on the plane: alpha = 2, nu_s = 1, nu_t = 1/2,
    c1 = Gamma(1/2) / (4 pi)^{3/2}: r_s = ln sqrt 8 - gs,
    r_t = gt + ln sqrt 4 - 2 gs, sigma_st = ln sqrt c1 - gst - gt/2 - gs."""
    gs, gt, gst = 0.4, -0.3, 0.9
    r_s, r_t, s_st = model_to_interpretable(gs, gt, gst, "plane")
    c1 = math.sqrt(math.pi) / (4 * math.pi) ** 1.5
    assert r_s == pytest.approx(0.5 * math.log(8) - gs, abs=1e-14)
    assert r_t == pytest.approx(gt + 0.5 * math.log(4) - 2 * gs, abs=1e-14)
    assert s_st == pytest.approx(0.5 * math.log(c1) - gst - 0.5 * gt - gs, abs=1e-14)


@pytest.mark.parametrize("nv", [2, 3])
def test_reference_coregional_table_matches_block_construction(nv):
    p = with_g3(coregional_problem(nv=nv, nt=3, nx=4, ny=3, nb_per=2, obs_per_node=2, seed=5))
    cm = ReferenceCoregionalCoefficients(nv, sigma_st=[0.1, -0.2, 0.3][:nv])
    rng = np.random.default_rng(nv)
    th = rng.normal(0, 0.3, cm.d)
    Q_tab = om.q_prior(p, cm.table(th))
    gam = lambda r_s, r_t, s: interpretable_to_model(r_s, r_t, s, "plane")
    Q_ref = om.coregional_q_prior_reference(p, th, cm.sigma_st, gam)
    d = abs(Q_tab - Q_ref).max()
    assert d <= 1e-12 * abs(Q_ref).max()
    # a planted mistake (lambda_0_2 and lambda_1_2 swapped) is visible
    if nv == 3:
        th2 = th.copy()
        th2[-2], th2[-1] = th[-1], th[-2]
        assert abs(om.q_prior(p, cm.table(th2)) - Q_ref).max() > 1e-6 * abs(Q_ref).max()


def test_f_with_pc_priors_is_the_exact_marginal_likelihood():
    """The identity of test_f_is_the_exact_marginal_likelihood with PC priors
    on every component of theta."""
    from pyinla_amd.coefficients import CoefficientModel
    from pyinla_amd.workloads import st_problem
    p = st_problem(nt=3, nx=4, ny=4, nb=2, seed=2)
    p.priors = HyperPriors([PenalizedComplexityPrior(0.01, 2.0, "r_s"), PenalizedComplexityPrior(0.05, 3.0, "r_t"),
                            PenalizedComplexityPrior(0.01, 1.0, "sigma_st"),
                            PenalizedComplexityPrior(0.01, 5.0, "prec_o")])
    cm = CoefficientModel("baseline")
    th = np.asarray(p.theta)
    f = om.evaluate_f(p, cm.table, th)
    Qp = om.q_prior(p, cm.table(th)).toarray()
    A = sps.csr_matrix(p.A).toarray()
    Sy = A @ np.linalg.inv(Qp) @ A.T + np.exp(-th[3]) * np.eye(A.shape[0])
    sign, ld = np.linalg.slogdet(Sy)
    logN = -0.5 * ld - 0.5 * p.y @ np.linalg.solve(Sy, p.y) - 0.5 * p.y.size * np.log(2 * np.pi)
    const = 0.5 * p.nb * np.log(p.fixed_prec) - 0.5 * p.y.size * np.log(2 * np.pi)
    assert abs(f + p.priors.log_prior(th) + logN - const) <= 1e-10 * abs(f)
