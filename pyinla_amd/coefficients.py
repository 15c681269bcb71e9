"""Coefficient tables c[m][k](theta) of Q_st = sum_{m,k} c[m][k] (M_m (x) S_k).

Rows m index the temporal matrices (M0, M1, M2), columns k the spatial
matrices (c0, g1, g2[, g3]).  Both parameterizations are written with torch
float64 scalars so that their Jacobians (needed by the adjoint gradient) come
from autograd on the CPU; they are a handful of scalar operations.
"""
from __future__ import annotations

import math

import numpy as np
import torch


def baseline_table(theta: torch.Tensor) -> torch.Tensor:
    """theta = (theta_s, theta_t, theta_e, ...):

    Q_st = e^{te} [ e^{2tt} M0 (x) (e^{4ts} c0 + 2 e^{2ts} g1 + g2)
                  + 2 e^{tt} M1 (x) (e^{2ts} c0 + g1) + M2 (x) c0 ]
    Returns a (3, 3) table over (M0, M1, M2) x (c0, g1, g2).
    """
    ts, tt, te = theta[0], theta[1], theta[2]
    z = torch.zeros((), dtype=torch.float64)
    one = torch.ones((), dtype=torch.float64)
    e = torch.exp(te)
    r0 = torch.stack([torch.exp(4 * ts), 2 * torch.exp(2 * ts), one]) * e * torch.exp(2 * tt)
    r1 = torch.stack([torch.exp(2 * ts), one, z]) * (2 * e * torch.exp(tt))
    r2 = torch.stack([one, z, z]) * e
    return torch.stack([r0, r1, r2])


def _gamma(theta, manifold: str):
    """(gamma_s, gamma_t, gamma_st) from (r_s, r_t, sigma_st).

    Restates SpatioTemporalSubModel._interpretable2compute
    (src/pyinla/submodels/spatio_temporal.py:151-197) with alpha_s = 2,
    alpha_t = 1, alpha_e = 1 and a 2-D spatial domain.
    """
    r_s, r_t, sigma_st = theta[0], theta[1], theta[2]
    alpha_s, alpha_t, alpha_e = 2, 1, 1
    alpha = alpha_e + alpha_s * (alpha_t - 0.5)
    nu_s = alpha - 1
    nu_t = alpha_t - 0.5
    gamma_s = 0.5 * math.log(8 * nu_s) - r_s
    gamma_t = r_t - 0.5 * math.log(8 * nu_t) + alpha_s * gamma_s
    if manifold == "sphere":
        cR_t = math.gamma(nu_t) / (math.gamma(alpha_t) * math.sqrt(4 * math.pi))
        c_s = 0.0
        for k in range(50):
            c_s = c_s + (2.0 * k + 1.0) / (4.0 * math.pi * (torch.exp(gamma_s) ** 2 + k * (k + 1)) ** alpha)
        gamma_st = 0.5 * math.log(cR_t) + 0.5 * torch.log(c_s) - 0.5 * gamma_t - sigma_st
    elif manifold == "plane":
        c1 = (math.gamma(nu_t) * math.gamma(nu_s)
              / (math.gamma(alpha_t) * math.gamma(alpha) * (4 * math.pi) ** 1.5))
        gamma_st = 0.5 * math.log(c1) - 0.5 * gamma_t - nu_s * gamma_s - sigma_st
    else:
        raise ValueError(f"Manifold not supported: {manifold}")
    return gamma_s, gamma_t, gamma_st


def reference_table(theta: torch.Tensor, manifold: str = "plane") -> torch.Tensor:
    """The reference parameterization (spatio_temporal.py:114-149):

    Q_st = e^{2gst} [ M0 (x) q3s + e^{gt} M1 (x) q2s + e^{2gt} M2 (x) q1s ],
    q1s = e^{2gs} c0 + g1, q2s = e^{4gs} c0 + 2 e^{2gs} g1 + g2,
    q3s = e^{6gs} c0 + 3 e^{4gs} g1 + 3 e^{2gs} g2 + g3.
    Returns a (3, 4) table over (M0, M1, M2) x (c0, g1, g2, g3).
    """
    gs, gt, gst = _gamma(theta, manifold)
    z = torch.zeros((), dtype=torch.float64)
    one = torch.ones((), dtype=torch.float64)
    es = torch.exp(gs)
    s = torch.exp(2 * gst)
    r0 = torch.stack([es ** 6, 3 * es ** 4, 3 * es ** 2, one]) * s
    r1 = torch.stack([es ** 4, 2 * es ** 2, one, z]) * (s * torch.exp(gt))
    r2 = torch.stack([es ** 2, one, z, z]) * (s * torch.exp(2 * gt))
    return torch.stack([r0, r1, r2])


class CoefficientModel:
    """Wraps a table function: value and Jacobian w.r.t. the first d_st thetas."""

    def __init__(self, kind: str = "baseline", manifold: str = "plane"):
        self.kind = kind
        self.manifold = manifold
        if kind == "baseline":
            self.fn = baseline_table
            self.nK = 3
        elif kind == "reference":
            self.fn = lambda th: reference_table(th, manifold)
            self.nK = 4
        else:
            raise ValueError(kind)
        self.nM = 3
        self.d_st = 3

    def table(self, theta) -> np.ndarray:
        th = torch.as_tensor(np.asarray(theta[: self.d_st], dtype=np.float64))
        return self.fn(th).detach().numpy().copy()

    def jacobian(self, theta) -> np.ndarray:
        """d table[m][k] / d theta_i as an array (d_st, nM, nK)."""
        th = torch.as_tensor(np.asarray(theta[: self.d_st], dtype=np.float64))
        J = torch.autograd.functional.jacobian(self.fn, th)  # (nM, nK, d_st)
        return J.permute(2, 0, 1).detach().numpy().copy()


# ---------------------------------------------------------------- priors

def gaussian_log_prior(theta, mean, prec):
    """sum_i -0.5 prec_i (theta_i - mean_i)^2 (prior_hyperparameters/gaussian.py:26-29)."""
    th = np.asarray(theta, dtype=np.float64)
    return float(np.sum(-0.5 * prec * (th - mean) ** 2))


def gaussian_log_prior_grad(theta, mean, prec):
    th = np.asarray(theta, dtype=np.float64)
    return -prec * (th - mean)


# ------------------------------------------------------------ coregional

def coregional_weights(lam: torch.Tensor, nv: int) -> torch.Tensor:
    """C[v][w][u]: Q_vw = sum_u C[v][w][u] Qu_u for the coregional precision.

    Restates CoregionalModel.construct_Q_prior (coregional_model.py:300-408)
    with the marginal scales sigma_u set to 1 (each field's own theta_e scales
    its Qu_u):
      nv = 2: q11 = Qu0 + l01^2 Qu1, q21 = -l01 Qu1, q22 = Qu1
      nv = 3: q11 = Qu0 + l01^2 Qu1 + l12^2 Qu2, q21 = -l01 Qu1 + l02 l12 Qu2,
              q31 = -l12 Qu2, q22 = Qu1 + l02^2 Qu2, q32 = -l02 Qu2, q33 = Qu2
    with lam = (l01,) or (l01, l02, l12); q_vw = q_wv.
    """
    z = torch.zeros((), dtype=torch.float64)
    one = torch.ones((), dtype=torch.float64)
    if nv == 2:
        l01 = lam[0]
        q00 = torch.stack([one, l01 ** 2])
        q10 = torch.stack([z, -l01])
        q11 = torch.stack([z, one])
        return torch.stack([torch.stack([q00, q10]), torch.stack([q10, q11])])
    if nv == 3:
        l01, l02, l12 = lam[0], lam[1], lam[2]
        q00 = torch.stack([one, l01 ** 2, l12 ** 2])
        q10 = torch.stack([z, -l01, l02 * l12])
        q20 = torch.stack([z, z, -l12])
        q11 = torch.stack([z, one, l02 ** 2])
        q21 = torch.stack([z, z, -l02])
        q22 = torch.stack([z, z, one])
        return torch.stack([torch.stack([q00, q10, q20]),
                            torch.stack([q10, q11, q21]),
                            torch.stack([q20, q21, q22])])
    raise ValueError("coregional models have 2 or 3 fields")


class CoregionalCoefficients:
    """Table (nv, nv, nM, nK) of the coregional precision.

    theta = (theta_s, theta_t, theta_e) per field, then the nv(nv-1)/2
    coregionalisation weights (l01[, l02, l12]), then one noise precision per
    field.  Field u's own precision uses the benchmark parameterization
    (baseline_table).
    """

    def __init__(self, nv: int):
        self.nv = nv
        self.nM, self.nK = 3, 3
        self.n_lam = nv * (nv - 1) // 2
        self.d_q = 3 * nv + self.n_lam

    def _fn(self, th: torch.Tensor) -> torch.Tensor:
        nv = self.nv
        tabs = torch.stack([baseline_table(th[3 * u:3 * u + 3]) for u in range(nv)])  # (u, m, k)
        C = coregional_weights(th[3 * nv:3 * nv + self.n_lam], nv)  # (v, w, u)
        return torch.einsum("vwu,umk->vwmk", C, tabs)

    def table(self, theta) -> np.ndarray:
        th = torch.as_tensor(np.asarray(theta[: self.d_q], dtype=np.float64))
        return self._fn(th).detach().numpy().copy()

    def jacobian(self, theta) -> np.ndarray:
        th = torch.as_tensor(np.asarray(theta[: self.d_q], dtype=np.float64))
        J = torch.autograd.functional.jacobian(self._fn, th)  # (v, w, m, k, d)
        return J.permute(4, 0, 1, 2, 3).detach().numpy().copy()
