"""Host-side optimizer driver (the reference's Optimizers, optimizers.py:21-166).

scipy.optimize.minimize drives fun/jac, which run on the GPU through the
wave-function classes.  RotoSolve is not provided by this engine.
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
            raise NotImplementedError("RotoSolve is not provided by the MI355X engine; use a gradient optimizer")
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
