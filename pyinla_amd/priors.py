"""Hyperparameter priors: log p(theta) and its gradient.

Restates the reference's two prior families on the hot path:

* Gaussian (src/pyinla/prior_hyperparameters/gaussian.py:26-29):
      log p = -1/2 precision (theta - mean)^2
* Penalized complexity (src/pyinla/prior_hyperparameters/penalized_complexity.py:28-85),
  with lambda from (alpha, u) per hyperparameter type (:28-43):
      r_s       (range, 2-D): lambda = -ln(alpha) u^{d/2},
                log p = ln(lambda) - lambda e^{-theta} - theta
      r_t       (temporal range): lambda = -ln(alpha) u^{1/2},
                log p = ln(lambda) - lambda e^{-theta/2} + ln(1/2) - theta/2
      sigma_st, sigma_e, prec_o: lambda = -ln(alpha) / u,
                log p = ln(lambda) - lambda e^{theta} + theta
  A PC prior built without a type (as CoregionalModel builds the priors of
  its sigma hyperparameters, coregional_model.py:160-164) has lambda = 0 and
  log p = 0 in the reference; it is kept that way here.

The gradients are the closed-form derivatives of these expressions; the
adjoint gradient of f subtracts them (f contains -log p(theta),
pyinla.py:1076-1110 via Model.evaluate_log_prior_hyperparameters,
model.py:455-478).
"""
from __future__ import annotations

import math

import numpy as np

PC_KINDS = ("r_s", "r_t", "sigma_st", "sigma_e", "prec_o")


class GaussianPrior:
    """N(mean, 1/precision) on one hyperparameter."""

    kind = "gaussian"

    def __init__(self, mean: float = 0.0, precision: float = 1.0):
        self.mean = float(mean)
        self.precision = float(precision)

    def log_prior(self, theta: float) -> float:
        return -0.5 * self.precision * (float(theta) - self.mean) ** 2

    def grad(self, theta: float) -> float:
        return -self.precision * (float(theta) - self.mean)


class PenalizedComplexityPrior:
    """Penalized-complexity prior of hyperparameter type ``kind`` with
    P(range < u) = alpha (r_s, r_t) or P(sigma / precision scale > u) = alpha
    (sigma_st, sigma_e, prec_o).  ``kind=None`` reproduces the untyped prior
    of the reference (log p = 0)."""

    def __init__(self, alpha: float, u: float, kind: str | None, spatial_dim: int = 2):
        if kind is not None and kind not in PC_KINDS:
            raise ValueError(f"unknown penalized-complexity hyperparameter type {kind!r}")
        if kind == "r_s" and spatial_dim != 2:
            raise ValueError("Not implemented for other than 2D spatial domains")
        self.alpha = float(alpha)
        self.u = float(u)
        self.kind = kind
        la = -math.log(self.alpha)
        if kind == "r_s":
            self.lam = la * self.u ** (0.5 * spatial_dim)
        elif kind == "r_t":
            self.lam = la * self.u ** 0.5
        elif kind in ("sigma_st", "sigma_e", "prec_o"):
            self.lam = la / self.u
        else:
            self.lam = 0.0

    def log_prior(self, theta: float) -> float:
        t, lam = float(theta), self.lam
        if self.kind == "r_s":
            return math.log(lam) - lam * math.exp(-t) - t
        if self.kind == "r_t":
            return math.log(lam) - lam * math.exp(-0.5 * t) + math.log(0.5) - 0.5 * t
        if self.kind in ("sigma_st", "sigma_e", "prec_o"):
            return math.log(lam) - lam * math.exp(t) + t
        return 0.0

    def grad(self, theta: float) -> float:
        t, lam = float(theta), self.lam
        if self.kind == "r_s":
            return lam * math.exp(-t) - 1.0
        if self.kind == "r_t":
            return 0.5 * lam * math.exp(-0.5 * t) - 0.5
        if self.kind in ("sigma_st", "sigma_e", "prec_o"):
            return -lam * math.exp(t) + 1.0
        return 0.0


class HyperPriors:
    """One prior per component of theta (the reference's
    Model.prior_hyperparameters list, evaluated component by component,
    model.py:455-478)."""

    def __init__(self, priors: list):
        self.priors = list(priors)

    @classmethod
    def gaussian(cls, mean, precision) -> "HyperPriors":
        mean = np.broadcast_to(np.asarray(mean, dtype=np.float64), np.shape(precision) or np.shape(mean))
        prec = np.broadcast_to(np.asarray(precision, dtype=np.float64), mean.shape)
        return cls([GaussianPrior(m, p) for m, p in zip(mean, prec)])

    def __len__(self) -> int:
        return len(self.priors)

    def log_prior(self, theta) -> float:
        th = np.asarray(theta, dtype=np.float64)
        if th.size != len(self.priors):
            raise ValueError(f"theta has {th.size} components, the priors {len(self.priors)}")
        return float(sum(p.log_prior(t) for p, t in zip(self.priors, th)))

    def grad(self, theta) -> np.ndarray:
        th = np.asarray(theta, dtype=np.float64)
        return np.array([p.grad(t) for p, t in zip(self.priors, th)], dtype=np.float64)


def priors_of(prob) -> HyperPriors:
    """The problem's priors: ``prob.priors`` if given, else independent
    Gaussians from ``prob.prior_mean`` / ``prob.prior_prec``."""
    pr = getattr(prob, "priors", None)
    if pr is not None:
        return pr if isinstance(pr, HyperPriors) else HyperPriors(pr)
    return HyperPriors.gaussian(prob.prior_mean, prob.prior_prec)
