"""ctypes wrapper of oracle/csrc/scan.c (test and CPU-baseline infrastructure only).

``propagate_state_c`` keeps the reference's structure: a Python loop over the
normal-ordered strings of each operator (operator_state_algebra.py:563-628)
calling one compiled scan per string (the Numba kernel there, C here).
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from .apply import string_masks
from .ci import CIInfo, parity_check

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "libsqoracle.so")


def build() -> str:
    """gcc -O3 with OpenMP; portable x86-64-v2 code (no -march=native)."""
    cmd = [
        "gcc", "-O3", "-march=x86-64-v2", "-fopenmp", "-fPIC", "-shared",
        os.path.join(HERE, "csrc", "scan.c"), "-o", SO + ".tmp",
    ]
    subprocess.run(cmd, check=True, capture_output=True)
    os.replace(SO + ".tmp", SO)
    return SO


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            build()
        L = ctypes.CDLL(SO)
        vp = ctypes.c_void_p
        L.oracle_map_build.argtypes = [vp, ctypes.c_int64, vp, vp, ctypes.c_int64]
        L.oracle_map_build.restype = None
        L.oracle_apply_string.argtypes = [
            vp, vp, vp, ctypes.c_int64, vp, vp, ctypes.c_int64, vp, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
            vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
        ]
        L.oracle_apply_string.restype = ctypes.c_int
        _lib = L
    return _lib


class HashedSpace:
    """CI space plus the C hash map det -> index (the reference's det2idx)."""

    def __init__(self, ci: CIInfo) -> None:
        self.ci = ci
        self.idx2det = np.ascontiguousarray(ci.idx2det, dtype=np.int64)
        n = len(self.idx2det)
        cap = 1
        while cap < 2 * n + 2:
            cap *= 2
        self.cap = cap
        self.keys = np.empty(cap, dtype=np.int64)
        self.vals = np.empty(cap, dtype=np.int64)
        lib().oracle_map_build(self.idx2det.ctypes.data, n, self.keys.ctypes.data, self.vals.ctypes.data, cap)
        self.pc = parity_check(ci.num_active_orbs)


def apply_string_c(hs: HashedSpace, state, key, factor, out, nthreads=1, do_unsafe=False):
    n_act = hs.ci.num_active_orbs
    a_string, cmask, amask = string_masks(key, n_act)
    a = np.asarray(a_string, dtype=np.int64)
    rc = lib().oracle_apply_string(
        state.ctypes.data, out.ctypes.data, hs.idx2det.ctypes.data, len(hs.idx2det), hs.keys.ctypes.data,
        hs.vals.ctypes.data, hs.cap, a.ctypes.data, len(a), cmask, amask, hs.pc.ctypes.data, 2 * n_act, float(factor),
        int(do_unsafe), int(nthreads),
    )
    if rc != 0:
        raise KeyError("operator maps outside the CI space")
    return out


def propagate_state_c(operators, state, hs: HashedSpace, nthreads=1, max_strings=None):
    """propagate_state without folding (operators already active-space)."""
    new = np.array(state, dtype=np.float64, copy=True)
    for op in operators[::-1]:
        tmp = np.zeros_like(new)
        items = list(op.operators.items())
        if max_strings is not None:
            items = items[:max_strings]
        for key, f in items:
            apply_string_c(hs, new, key, f, tmp, nthreads)
        new = tmp
    return new
