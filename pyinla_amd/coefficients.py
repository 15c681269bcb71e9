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


def interpretable_to_model(r_s, r_t, sigma_st, manifold: str = "plane"):
    """(gamma_s, gamma_t, gamma_st) of the model scale from the interpretable
    (r_s, r_t, sigma_st): SpatioTemporalSubModel._interpretable2compute
    (spatio_temporal.py:151-197), 2-D domain, alpha_s = 2, alpha_t = 1,
    alpha_e = 1.  Accepts floats or torch scalars."""
    th = torch.stack([torch.as_tensor(r_s, dtype=torch.float64), torch.as_tensor(r_t, dtype=torch.float64),
                      torch.as_tensor(sigma_st, dtype=torch.float64)])
    gs, gt, gst = _gamma(th, manifold)
    return gs, gt, gst


def model_to_interpretable(gamma_s, gamma_t, gamma_st, manifold: str = "plane", dim_spatial_domain: int = 2):
    """(r_s, r_t, sigma_st) from the model scale (gamma_s, gamma_t, gamma_st):
    SpatioTemporalSubModel.convert_theta_from_model2interpret
    (spatio_temporal.py:199-263), the inverse of interpretable_to_model:
      r_s = ln(sqrt(8 nu_s) / e^{gs}),  r_t = ln(e^{gt} sqrt(8 nu_t) / e^{alpha_s gs}),
      plane:  sigma_st = ln(sqrt(c1) / (e^{gst} sqrt(e^{gt}) e^{gs (alpha - d/2)})),
              c1 = Gamma(nu_t) Gamma(nu_s) / (Gamma(alpha_t) Gamma(alpha) (4 pi)^{d/2} (4 pi)^{1/2})
      sphere: sigma_st = ln(sqrt(c_rt c_s) / (e^{gst} sqrt(e^{gt}))),
              c_rt = Gamma(alpha_t - 1/2) / (Gamma(alpha_t) sqrt(4 pi)),
              c_s = sum_{k<50} (2k+1) / (4 pi (e^{2 gs} + k(k+1))^alpha)."""
    if dim_spatial_domain != 2:
        raise ValueError("Only 2D spatial domain is supported for now.")
    alpha_t, alpha_s, alpha_e = 1, 2, 1
    alpha = alpha_e + alpha_s * (alpha_t - 0.5)
    nu_s = alpha - 1
    nu_t = alpha_t - 0.5
    gs, gt, gst = float(gamma_s), float(gamma_t), float(gamma_st)
    r_s = math.log(math.sqrt(8 * nu_s) / math.exp(gs))
    r_t = math.log(math.exp(gt) * math.sqrt(8 * nu_t) / math.exp(gs) ** alpha_s)
    if manifold == "sphere":
        c_rt = math.gamma(alpha_t - 0.5) / (math.gamma(alpha_t) * math.sqrt(4 * math.pi))
        c_s = 0.0
        for k in range(50):
            c_s += (2.0 * k + 1) / (4 * math.pi * (math.exp(gs) ** 2 + k * (k + 1)) ** alpha)
        sigma_st = math.log(math.sqrt(c_rt * c_s) / (math.exp(gst) * math.sqrt(math.exp(gt))))
    elif manifold == "plane":
        c1_scale = (4 * math.pi) ** (dim_spatial_domain / 2.0) * (4 * math.pi) ** 0.5
        c1 = math.gamma(nu_t) * math.gamma(nu_s) / (math.gamma(alpha_t) * math.gamma(alpha) * c1_scale)
        sigma_st = math.log(math.sqrt(c1) / ((math.exp(gst) * math.sqrt(math.exp(gt)))
                                             * math.exp(gs) ** (alpha - dim_spatial_domain / 2)))
    else:
        raise ValueError(f"Manifold not supported: {manifold}")
    return r_s, r_t, sigma_st


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

    def noise_index(self, v: int = 0) -> int:
        """theta index of the noise precision of observation set v."""
        return self.d_st + v

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

    def noise_index(self, v: int) -> int:
        return self.d_q + v

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


class ReferenceCoregionalCoefficients:
    """The reference's own coregional parameterization
    (CoregionalModel.construct_Q_prior, coregional_model.py:274-408, with the
    theta layout of coregional_model.py:96-150):

      theta = (r_s, r_t, prec_o) per field v, then sigma_0 .. sigma_{nv-1},
              then lambda_0_1 [, lambda_0_2, lambda_1_2];
      Qu_v   = the reference ST precision of field v (reference_table) at
               (r_s_v, r_t_v) and the field's fixed sigma_st (its sigma_st
               hyperparameter is taken out of theta, :104-122);
      Q_vw   = sum_u C[v][w][u](lambda) Qu_u / sigma_u^2, sigma_u = e^{theta_sigma_u},
    with the weights of coregional_weights (q11 = Qu0/s0^2 + l01^2 Qu1/s1^2
    + l12^2 Qu2/s2^2, ...).  Table (nv, nv, 3, 4) over (M0, M1, M2) x
    (c0, g1, g2, g3); the Jacobian covers every theta component (zero rows
    for the noise precisions)."""

    def __init__(self, nv: int, manifold: str = "plane", sigma_st=0.0):
        if nv not in (2, 3):
            raise ValueError("coregional models have 2 or 3 fields")
        self.nv = nv
        self.manifold = manifold
        self.sigma_st = np.broadcast_to(np.asarray(sigma_st, dtype=np.float64), (nv,)).copy()
        self.nM, self.nK = 3, 4
        self.n_lam = nv * (nv - 1) // 2
        self.d = 3 * nv + nv + self.n_lam
        self.d_q = self.d  # the table depends on theta up to here (noise rows are zero)

    def noise_index(self, v: int) -> int:
        return 3 * v + 2

    def sigma_index(self, u: int) -> int:
        return 3 * self.nv + u

    def _fn(self, th: torch.Tensor) -> torch.Tensor:
        nv = self.nv
        tabs = []
        for u in range(nv):
            t_u = torch.stack([th[3 * u], th[3 * u + 1],
                               torch.as_tensor(self.sigma_st[u], dtype=torch.float64)])
            sig = torch.exp(th[3 * nv + u])
            tabs.append(reference_table(t_u, self.manifold) / sig ** 2)
        tabs = torch.stack(tabs)  # (u, m, k)
        C = coregional_weights(th[4 * nv:4 * nv + self.n_lam], nv)  # (v, w, u)
        return torch.einsum("vwu,umk->vwmk", C, tabs)

    def table(self, theta) -> np.ndarray:
        th = torch.as_tensor(np.asarray(theta[: self.d], dtype=np.float64))
        return self._fn(th).detach().numpy().copy()

    def jacobian(self, theta) -> np.ndarray:
        th = torch.as_tensor(np.asarray(theta[: self.d], dtype=np.float64))
        J = torch.autograd.functional.jacobian(self._fn, th)  # (v, w, m, k, d)
        return J.permute(4, 0, 1, 2, 3).detach().numpy().copy()
