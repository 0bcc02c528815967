"""Host-side optimizer driver (the reference's Optimizers, optimizers.py:21-166)
and RotoSolve (optimizers.py:169-453).

scipy.optimize.minimize drives fun/jac, which run on the GPU through the
wave-function classes.  RotoSolve asks the wave function for the energies at
2R+1 values of one parameter; the UPS class computes them in one batched
engine launch.
"""

from __future__ import annotations

import time
from collections.abc import Callable, Sequence
from functools import partial
from typing import Any

import numpy as np
import scipy.optimize


class Result:
    """Optimization result (x, fun, success)."""

    def __init__(self) -> None:
        self.x: np.ndarray
        self.fun: float
        self.success: bool


class Optimizers:
    def __init__(
        self,
        fun: Callable,
        method: str,
        grad: Callable | None = None,
        maxiter: int = 1000,
        tol: float = 10e-8,
        is_silent: bool = False,
        energy_eval_callback: Callable[[], int] | None = None,
        std_callback: Callable[[], float] | None = None,
    ) -> None:
        self.fun = fun
        self.grad = grad
        self.method = method.lower()
        self.maxiter = maxiter
        self.tol = tol
        self.is_silent = is_silent
        self.energy_eval_callback = energy_eval_callback
        self.std_callback = std_callback
        self._start = 0.0
        self._iteration = 0

    def _print_progress(self, x: Sequence[float], fun: Callable, silent: bool = False) -> None:
        if silent:
            return
        e = fun(list(x))
        e_str = f"{np.mean(e):3.16f}" if isinstance(e, np.ndarray) else f"{e:3.16f}"
        time_str = f"{time.time() - self._start:7.2f}"
        evals = str(self.energy_eval_callback()) if self.energy_eval_callback else ""
        std = ""
        if self.std_callback is not None:
            var = self.std_callback()
            if var is not None:
                std = f" | {np.sqrt(var):.6e}"
        print(f"--------{str(self._iteration + 1).center(11)} | {time_str.center(18)} | {e_str.center(27)} | {evals.center(20)}{std}")
        self._iteration += 1
        self._start = time.time()

    def minimize(self, x0: Sequence[float], extra_options: dict[str, Any] | None = None) -> Result:
        self._start = time.time()
        self._iteration = 0
        cb = partial(self._print_progress, fun=self.fun, silent=self.is_silent)
        opts = {"maxiter": self.maxiter, "disp": True}
        if self.method in ("bfgs", "l-bfgs-b", "slsqp"):
            kw = {"jac": self.grad} if self.grad is not None else {}
            res = scipy.optimize.minimize(self.fun, x0, method=self.method, tol=self.tol, callback=cb, options=opts, **kw)
        elif self.method in ("cobyla", "cobyqa"):
            res = scipy.optimize.minimize(self.fun, x0, method=self.method, tol=self.tol, callback=cb, options=opts)
        elif self.method == "rotosolve":
            if not isinstance(extra_options, dict):
                raise TypeError("extra_options is not set, but is required for RotoSolve")
            if "R" not in extra_options:
                raise ValueError(f"Expected option 'R' in extra_options, got {extra_options.keys()}")
            if "param_names" not in extra_options:
                raise ValueError(f"Expected option 'param_names' in extra_options, got {extra_options.keys()}")
            roto = RotoSolve(extra_options["R"], extra_options["param_names"], maxiter=self.maxiter, tol=self.tol,
                             callback=cb)
            res = roto.minimize(self.fun, x0, f_rotosolve_optimized=extra_options.get("f_rotosolve_optimized"))
        else:
            raise ValueError(f"Got an unkonwn optimizer {self.method}")
        out = Result()
        out.x = res.x
        out.fun = res.fun
        out.success = res.success
        if not out.success:
            print("Optimization failed.")
            if hasattr(res, "message"):
                print(res.message)
        return out


# ---------------------------------------------------------------------------
# RotoSolve (optimizers.py:169-453).  For a generator whose eigenvalue
# differences take R distinct positive values, E(x) in one parameter x is a
# trigonometric polynomial of order R, fixed by its values at
# x_mu = 2 mu pi / (2R+1), mu = -R..R (Dirichlet-kernel interpolation):
#   E(x) = sum_mu E(x_mu) sinc((2R+1)(x - x_mu) / (2 pi)) / sinc((x - x_mu) / (2 pi))
# with numpy's normalised sinc.


def shift_points(R: int) -> list[float]:
    """x_mu = 2 mu pi / (2R+1) for mu = -R..R (optimizers.py:289-292, 381-383)."""
    return [2 * mu / (2 * R + 1) * np.pi for mu in range(-R, R + 1)]


def _per_state(energy_vals) -> tuple[np.ndarray, bool]:
    """(energies as [2R+1, n_states], state-averaged?)."""
    if isinstance(energy_vals[0], (float, np.floating)):
        return np.asarray(energy_vals, dtype=np.float64)[:, None], False
    return np.asarray([np.asarray(e, dtype=np.float64) for e in energy_vals]), True


def reconstructed_f(x_vals, energy_vals, R: int) -> np.ndarray:
    """E(x) rebuilt from the 2R+1 samples (optimizers.py:321-368).  For lists of
    per-state energies the state sum is divided by the number of samples, as
    the reference does (optimizers.py:367)."""
    x = np.atleast_1d(np.asarray(x_vals, dtype=np.float64))
    ev, averaged = _per_state(energy_vals)
    d = x[None, :] - np.asarray(shift_points(R))[:, None]  # [mu, point]
    w = np.sinc((2 * R + 1) / 2 * d / np.pi) / np.sinc(0.5 * d / np.pi)
    e = (ev.sum(axis=1)[:, None] * w).sum(axis=0)
    return e / len(energy_vals) if averaged else e


def _sinc_derivative(u: np.ndarray) -> np.ndarray:
    """d/du of numpy's sinc(u) = sin(pi u)/(pi u); 0 at u = 0 (optimizers.py:390-406)."""
    u = np.asarray(u, dtype=np.float64)
    du = np.zeros_like(u)
    nz = np.abs(u) > 1e-12
    un = u[nz]
    du[nz] = np.cos(np.pi * un) / un - np.sin(np.pi * un) / (np.pi * un**2)
    return du


def reconstructed_f_derivative(x_vals, energy_vals, R: int) -> np.ndarray:
    """dE/dx of reconstructed_f by the quotient rule (optimizers.py:409-497)."""
    x = np.atleast_1d(np.asarray(x_vals, dtype=np.float64))
    ev, averaged = _per_state(energy_vals)
    A, B = (2 * R + 1) / 2.0, 0.5
    d = x[None, :] - np.asarray(shift_points(R))[:, None]
    u, v = A * d / np.pi, B * d / np.pi
    s1, s2 = np.sinc(u), np.sinc(v)
    s1p = _sinc_derivative(u) * (A / np.pi)
    s2p = _sinc_derivative(v) * (B / np.pi)
    de = (ev.sum(axis=1)[:, None] * (s1p * s2 - s1 * s2p) / s2**2).sum(axis=0)
    return de / len(energy_vals) if averaged else de


def get_energy_evals(f: Callable, x, idx: int, R: int) -> list:
    """f at x with x[idx] = x_mu for every shift point (optimizers.py:299-318)."""
    x = list(x).copy()
    out = []
    for x_mu in shift_points(R):
        x[idx] = x_mu
        out.append(f(x))
    return out


def get_energy_evals_optimized(f: Callable, x, idx: int, R: int) -> list:
    """f(x, shift points, idx): all shifted energies in one call (optimizers.py:371-387)."""
    return f(x, shift_points(R), idx)


def _wrap(t: float) -> float:
    """Into (-pi, pi] by whole turns, as optimizers.py:258-261."""
    while t < np.pi:
        t += 2 * np.pi
    while t > np.pi:
        t -= 2 * np.pi
    return t


class RotoSolve:
    """Parameter-by-parameter exact minimisation (optimizers.py:169-296).

    Per parameter: the 2R+1 shifted energies rebuild E(x) exactly; its minimum
    is located on a 10^4-point grid over [-pi, pi] and refined with BFGS
    (tol 1e-12).  A sweep over all parameters is one iteration; the run stops
    when the energy after a sweep changes by less than tol, or after six
    sweeps that raised it.
    """

    def __init__(self, R: dict[str, int], param_names: Sequence[str], maxiter: int = 1000, tol: float = 1e-8,
                 callback: Callable[[list[float]], None] | None = None) -> None:
        self._callback = callback
        self.max_iterations = maxiter
        self.threshold = tol
        self.max_fail = 6
        self._R = R
        self._param_names = param_names

    def minimize(self, f: Callable, x0: Sequence[float], f_rotosolve_optimized: Callable | None = None) -> Result:
        f_best = float(10**20)
        x = list(x0).copy()
        x_best = x.copy()
        fails = 0
        success = False
        grid = np.linspace(-np.pi, np.pi, int(1e4))
        for _ in range(self.max_iterations):
            for i, name in enumerate(self._param_names):
                R = self._R[name]
                if f_rotosolve_optimized is not None:
                    e_vals = get_energy_evals_optimized(f_rotosolve_optimized, x, i, R)
                else:
                    e_vals = get_energy_evals(f, x, i, R)
                theta = grid[np.argmin(reconstructed_f(grid, e_vals, R))]
                res = scipy.optimize.minimize(
                    lambda t, ev=e_vals, r=R: float(reconstructed_f(t, ev, r)[0]),
                    x0=[theta],
                    jac=lambda t, ev=e_vals, r=R: reconstructed_f_derivative(t, ev, r),
                    method="BFGS",
                    tol=1e-12,
                )
                x[i] = _wrap(float(res.x[0]))
            f_tmp = f(x)
            f_new = float(np.mean(f_tmp)) if isinstance(f_tmp, np.ndarray) else f_tmp
            if self._callback is not None:
                self._callback(x)
            if abs(f_best - f_new) < self.threshold:
                f_best = f_new
                x_best = x.copy()
                success = True
                break
            if (f_new - f_best) > 0.0:
                fails += 1
            else:
                f_best = f_new
                x_best = x.copy()
            if fails == self.max_fail:
                print("Three energy raises detected.")
                break
        out = Result()
        out.x = np.array(x_best)
        out.fun = f_best
        out.success = success
        return out
