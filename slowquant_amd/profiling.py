"""Per-kernel timing with HIP events recorded by the engine on its stream."""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native as N

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
FP64_MFMA_PEAK_TFLOPS = 78.6  # MI355X dense FP64 matrix
MFMA_CLASSES = ("sigma_gemm_mfma", "sigma_rows_mfma", "alpha_gemm_mfma", "rdm_gram_mfma", "orbital_gemm_mfma")


def collect() -> dict:
    L = N.lib()
    n = L.sq_profile_nclasses()
    ms = np.zeros(n)
    cnt = np.zeros(n, dtype=np.int64)
    by = np.zeros(n)
    fl = np.zeros(n)
    N.check(L.sq_profile_collect(n, ms.ctypes.data, cnt.ctypes.data, by.ctypes.data, fl.ctypes.data))
    out = {}
    for c in range(n):
        if cnt[c] == 0:
            continue
        name = L.sq_profile_class_name(c).decode()
        avg_ms = ms[c] / cnt[c]
        rec = {
            "launches": int(cnt[c]),
            "total_ms": float(ms[c]),
            "avg_us": 1e3 * float(avg_ms),
            "bytes_per_launch": float(by[c] / cnt[c]),
            "flops_per_launch": float(fl[c] / cnt[c]),
        }
        secs = ms[c] / 1e3
        rec["achieved_GBs"] = float(by[c] / secs / 1e9) if secs > 0 else 0.0
        rec["achieved_TFLOPs"] = float(fl[c] / secs / 1e12) if secs > 0 else 0.0
        out[name] = rec
    return out


def profile(fn, reps: int = 3) -> dict:
    """Run fn() reps times with event bracketing of every engine launch."""
    L = N.lib()
    torch.cuda.synchronize()
    N.check(L.sq_profile_enable(1))
    try:
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        kernels = collect()
    finally:
        N.check(L.sq_profile_enable(0))
    for k in kernels.values():
        k["launches_per_call"] = k["launches"] / reps
        k["ms_per_call"] = k["total_ms"] / reps
    return kernels


def class_index(name: str) -> int:
    L = N.lib()
    for c in range(L.sq_profile_nclasses()):
        if L.sq_profile_class_name(c).decode() == name:
            return c
    raise ValueError(f"unknown kernel class {name}")


def timed_profile(fn, steps: int, warmup: int, classes: tuple[str, ...] | None = None) -> tuple[float, dict]:
    """Seconds per call of fn() over ``steps`` calls after ``warmup`` untimed
    ones, with the launches of the kernel ``classes`` (all when None) in the
    timed calls bracketed by HIP events on their own stream; returns
    (seconds per call, per-class kernel table of the timed classes)."""
    import time

    L = N.lib()
    mask = 0xFFFFFFFF if classes is None else sum(1 << class_index(c) for c in classes)
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    N.check(L.sq_profile_select(mask))
    N.check(L.sq_profile_enable(1))
    try:
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        secs = (time.perf_counter() - t0) / steps
        kernels = collect()
    finally:
        N.check(L.sq_profile_enable(0))
        N.check(L.sq_profile_select(0xFFFFFFFF))
    for k in kernels.values():
        k["launches_per_call"] = k["launches"] / steps
        k["ms_per_call"] = k["total_ms"] / steps
    return secs, kernels


def dominant(kernels: dict) -> dict:
    name = max(kernels, key=lambda k: kernels[k]["total_ms"])
    k = kernels[name]
    if name in MFMA_CLASSES:
        return {"name": name, "bound": "mfma", "achieved": k["achieved_TFLOPs"], "peak": FP64_MFMA_PEAK_TFLOPS,
                "unit": "TFLOP/s", "avg_us": k["avg_us"], "per_launch": k["flops_per_launch"]}
    return {"name": name, "bound": "hbm", "achieved": k["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "avg_us": k["avg_us"], "per_launch": k["bytes_per_launch"]}


def profile_energy_grad(space, prog, H, theta, psi0, reps: int = 3, method: str = "adjoint") -> dict:
    kernels = profile(lambda: prog.energy_grad(H, theta, psi0, method), reps)
    return {"kernels": kernels, "dominant": dominant(kernels)}
