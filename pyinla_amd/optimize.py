"""Mode search over theta with scipy L-BFGS-B, as PyINLA.minimize does.

Restates src/pyinla/core/pyinla.py:330-510: scipy.optimize.minimize with
method "L-BFGS-B", jac=True, the options of BFGSConfig
(configs/pyinla_config.py:19-30, ftol 1e-22) and the callback's early exits:
after more than ``f_reduction_lag`` accepted iterations, stop when
|f_values[-lag] - f| < f_reduction_tol; likewise for ||theta_values[-lag] -
theta|| < theta_reduction_tol.  The objective returns (f, grad) from the
device: the adjoint gradient ("adjoint"), the reference's sequential FD loop
("fd") or the same FD with all factorizations batched ("fd_batched").
"""
from __future__ import annotations

import time

import numpy as np
from scipy import optimize


class _EarlyExit(Exception):
    pass


def minimize(model, theta0, gradient: str = "adjoint", max_iter: int = 100, maxcor: int = 10, maxls: int = 20,
             gtol: float = 1e-1, f_reduction_lag: int = 3, f_reduction_tol: float = 1e-4,
             theta_reduction_lag: int = 3, theta_reduction_tol: float = 1e-4, eps_gradient_f: float = 1e-3,
             fd_wave: int | None = None, verbose: bool = False) -> dict:
    if gradient not in ("adjoint", "fd", "fd_batched"):
        raise ValueError(gradient)
    hist = {"f_values": [], "theta_values": [], "objective_time": [], "n_obj": 0}
    last = {}

    def objective(theta):
        tic = time.perf_counter()
        if gradient == "adjoint":
            f, g = model.value_and_grad(theta)
        elif gradient == "fd":
            f, g = model.fd_gradient(theta, eps=eps_gradient_f)
        else:
            f, g = model.fd_gradient_batched(theta, eps=eps_gradient_f, wave=fd_wave)
        hist["objective_time"].append(time.perf_counter() - tic)
        hist["n_obj"] += 1
        if verbose:
            print(f"  objective call {hist['n_obj']:3d} ({gradient}): f={float(f):.8f} "
                  f"in {hist['objective_time'][-1]:.2f} s", flush=True)
        last["grad"] = np.asarray(g, dtype=np.float64).copy()
        return float(f), np.asarray(g, dtype=np.float64)

    state = {"accepted": 0, "result": None}

    def callback(intermediate_result):
        theta_i = np.array(intermediate_result.x, copy=True)
        fun_i = float(intermediate_result.fun)
        state["accepted"] += 1
        hist["theta_values"].append(theta_i)
        hist["f_values"].append(fun_i)
        if verbose:
            print(f"iter {state['accepted']:3d} f={fun_i:.8f} |g|={np.linalg.norm(last.get('grad', 0)):.3e}",
                  flush=True)
        k = state["accepted"]
        if k > f_reduction_lag and abs(hist["f_values"][-f_reduction_lag] - fun_i) < f_reduction_tol:
            state["result"] = (theta_i, fun_i, "f reduction below tolerance")
            raise _EarlyExit()
        if k > theta_reduction_lag and np.linalg.norm(hist["theta_values"][-theta_reduction_lag] - theta_i) \
                < theta_reduction_tol:
            state["result"] = (theta_i, fun_i, "theta change below tolerance")
            raise _EarlyExit()

    tic = time.perf_counter()
    try:
        res = optimize.minimize(objective, np.asarray(theta0, dtype=np.float64), method="L-BFGS-B", jac=True,
                                options={"maxiter": max_iter, "maxcor": maxcor, "maxls": maxls, "gtol": gtol,
                                         "disp": False, "ftol": 1e-22},
                                callback=callback)
        theta, fval, msg = res.x, float(res.fun), str(res.message)
    except _EarlyExit:
        theta, fval, msg = state["result"]
    return {"theta": np.asarray(theta), "f": fval, "message": msg, "iterations": state["accepted"],
            "objective_calls": hist["n_obj"], "f_values": hist["f_values"], "theta_values": hist["theta_values"],
            "objective_time": hist["objective_time"], "wall_time": time.perf_counter() - tic,
            "grad_f": last.get("grad")}
