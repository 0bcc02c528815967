"""Kernel-shape autotuner with an on-disk cache.

The reference's only tuning knob is its Numba thread count
(operator_state_algebra.py:23-30); the engine's kernels have shapes instead,
chosen here by timing on first use and remembered per device and workload:

* single instance (a space and a gate program): the fused rotation stages'
  tile size (sq_program_set_tile_shape: elements and threads per tile) by timing whole energy + gradient
  evaluations, then the row sigma's rows per block and alpha-GEMM split
  (sq_sigma_configure) by timing sigma alone;
* batched instances: the persistent launch's workgroup count and kernel shape
  (engine.autotune_batch), keyed by the instance list.

Every choice computes the same quantities (tests check them against the
default shapes to 1e-12); only the time differs.  The cache is a JSON file,
$SQ_AUTOTUNE_CACHE or ~/.cache/slowquant_amd/autotune.json, written
atomically; a missing or unreadable file means "tune again".
"""

from __future__ import annotations

import hashlib
import json
import os
import tempfile
import time

import numpy as np
import torch

from . import _native as N
from . import engine as E

TILE_TARGETS = (0, 256, 512, 1024, 2048, 4096)
TILE_THREADS = (512, 256)
SIGMA_ROWS = (0, 1, 2, 3, 4)
SIGMA_SPLITS = (1, 2, 4)


def cache_path() -> str:
    p = os.environ.get("SQ_AUTOTUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "slowquant_amd", "autotune.json")


def _load() -> dict:
    try:
        with open(cache_path()) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def _store(key: str, value: dict) -> None:
    path = cache_path()
    d = _load()
    d[key] = value
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=os.path.dirname(path) or ".", prefix=".autotune.")
    with os.fdopen(fd, "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _device_key() -> str:
    p = torch.cuda.get_device_properties(0)
    arch = getattr(p, "gcnArchName", "") or "gpu"
    return f"{arch}:{p.multi_processor_count}cu|{N.lib().sq_version().decode()}"


def program_key(space: E.Space, prog: E.Program) -> str:
    h = hashlib.sha1(json.dumps([prog.types, prog.indices]).encode()).hexdigest()[:16]
    return f"{_device_key()}|ups|{space.n_orb},{space.n_alpha},{space.n_beta}|{h}"


def _time(fn, reps: int) -> float:
    fn()  # first call builds plans and tables outside the timing
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def apply(space: E.Space, prog: E.Program, choice: dict) -> None:
    prog.set_tile_shape(int(choice.get("tile_target", 0)), int(choice.get("tile_threads", 0)))
    space.configure_sigma(int(choice.get("rows_per_block", 0)), int(choice.get("nsplit", 0)))


def tune(space: E.Space, prog: E.Program, H: E.Hamiltonian, theta, psi0: torch.Tensor, reps: int = 5,
         use_cache: bool = True) -> dict:
    """Pick the tile shape and the sigma shape for (space, prog) by timing,
    apply them, and return {"tile_target", "tile_threads", "rows_per_block",
    "nsplit", "ms_per_eval", "cached"}."""
    key = program_key(space, prog)
    if use_cache:
        hit = _load().get(key)
        if isinstance(hit, dict):
            try:
                apply(space, prog, hit)
                return {**hit, "cached": True}
            except (ValueError, RuntimeError):
                pass  # a stale entry that no longer fits: tune again
    th = np.asarray(theta, dtype=np.float64)

    def evaluate():
        prog.energy_grad(H, th, psi0, "adjoint")

    best_tile, best_t = (0, 0), float("inf")
    for nt in TILE_THREADS:
        for t in TILE_TARGETS:
            try:
                prog.set_tile_shape(t, nt)
                dt = _time(evaluate, reps)
            except (ValueError, RuntimeError):
                continue  # a tile size the stages cannot use
            if dt < best_t:
                best_tile, best_t = (t, nt), dt
    prog.set_tile_shape(*best_tile)
    psi = prog.state_prep(th, psi0)
    best_sig, best_st = (0, 0), float("inf")
    for r in SIGMA_ROWS:
        for ns in SIGMA_SPLITS:
            try:
                space.configure_sigma(r, ns)
            except ValueError:
                continue
            dt = _time(lambda: H.sigma(psi, want_energy=False), 4 * reps)
            if dt < best_st:
                best_sig, best_st = (r, ns), dt
    space.configure_sigma(*best_sig)
    ms = 1e3 * _time(evaluate, reps)
    choice = {"tile_target": best_tile[0], "tile_threads": best_tile[1], "rows_per_block": best_sig[0],
              "nsplit": best_sig[1], "ms_per_eval": ms}
    if use_cache:
        _store(key, choice)
    return {**choice, "cached": False}


def batch_key(instances) -> str:
    desc = [(s.n_orb, s.n_alpha, s.n_beta, p.types, p.indices) for s, p, _ in instances]
    h = hashlib.sha1(json.dumps(desc).encode()).hexdigest()[:16]
    return f"{_device_key()}|batch|{len(instances)}|{h}"


def tune_batch(instances, theta, use_cache: bool = True, **kw) -> E.Batch:
    """engine.autotune_batch with its choice (workgroups, kernel shape) cached."""
    key = batch_key(instances)
    if use_cache:
        hit = _load().get(key)
        if isinstance(hit, dict):
            ts = hit.get("tile_shape")
            if ts:
                for p in E.grouped_programs(instances):
                    p.set_tile_shape(*ts)
            b = E.Batch(instances, int(hit["workgroups"]))
            b.tile_shape = tuple(ts) if ts else None
            b.set_theta(theta)
            b.configure(*hit["shape"])
            b.tuned_ms = hit.get("ms")
            b.cached = True
            return b
    b = E.autotune_batch(instances, theta, **kw)
    b.cached = False
    if use_cache:
        _store(key, {"workgroups": b.n_workgroups, "shape": list(b.shape), "ms": b.tuned_ms,
                     "tile_shape": list(b.tile_shape) if b.tile_shape else None})
    return b
