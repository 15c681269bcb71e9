"""The reference's end-to-end pipeline on the device-resident model.

run() follows PyINLA.run (src/pyinla/core/pyinla.py:291-328):
  1. mode of the hyperparameters theta (PyINLA.minimize, pyinla.py:330-510;
     here optimize.minimize with the adjoint or finite-difference gradient);
  2. covariance of theta at the mode from the Hessian of f
     (PyINLA.compute_covariance_hp, pyinla.py:747-910);
  3. marginal variances of the latent parameters and of the observations
     (pyinla.py:912-1023).
The result dictionary has the reference's keys.
"""
from __future__ import annotations

import numpy as np

from .optimize import minimize


def run(model, theta0, gradient: str = "adjoint", hessian: str = "fd", eps_hessian_f: float = 5e-3,
        **minimize_kwargs) -> dict:
    """Mode, cov(theta) and marginal variances; ``gradient`` is "adjoint",
    "fd" or "fd_batched" (the objective of the minimization), ``hessian``
    "fd" (the reference's stencil, eps_hessian_f = 5e-3 as
    pyinla_config.py:52) or "adjoint" (differences of adjoint gradients)."""
    res = minimize(model, theta0, gradient=gradient, **minimize_kwargs)
    theta_star = np.asarray(res["theta"], dtype=np.float64)
    model.forward(theta_star)  # x* at the mode (the reference's model.x)
    x_star = model.x.detach().cpu().numpy().copy()
    cov_theta = model.cov_theta(theta_star, eps=eps_hessian_f, method=hessian)
    lat, obs = model.marginal_variances(theta_star)
    return {
        "theta": theta_star,
        "x": x_star,
        "f": float(res["f"]),
        "grad_f": res["grad_f"],
        "f_values": res["f_values"],
        "theta_values": res["theta_values"],
        "cov_theta": cov_theta,
        "marginal_variances_latent": lat.detach().cpu().numpy(),
        "marginal_variances_observations": np.concatenate([o.detach().cpu().numpy() for o in obs]),
        "message": res["message"],
        "iterations": res["iterations"],
    }
