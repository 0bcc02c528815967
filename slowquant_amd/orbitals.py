"""Orbital rotations, AO -> MO transforms and the orbital gradient on the device.

Device counterparts of c_mo = C expm(-kappa) (ups_wavefunction.py:367-385),
one_/two_electron_integral_transform (integralfunctions.py:186-211) and
get_orbital_gradient (density_matrix.py:181-230).  Every contraction runs as
an FP64-MFMA GEMM in libsqhip.so (sq_orbital_rotation, sq_ao2mo,
sq_orbital_gradient); matrices are float64 CUDA tensors, row-major.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native as N
from .engine import _ptr, _stream, require_gpu

_WORK: dict[tuple, torch.Tensor] = {}


def _work(n_ao: int, n_mo: int, n_inactive: int = 0, n_active: int = 1, n_kappa: int = 0) -> torch.Tensor:
    key = (n_ao, n_mo, n_inactive, n_active, n_kappa)
    if key not in _WORK:
        nbytes = int(N.lib().sq_orbital_work_bytes(n_ao, n_mo, n_inactive, n_active, n_kappa))
        if nbytes < 0:
            raise ValueError("invalid orbital dimensions")
        _WORK.clear()  # one workspace at a time: shapes change rarely
        _WORK[key] = torch.empty(max(nbytes // 8, 1), dtype=torch.float64, device="cuda")
    return _WORK[key]


def _dev(x) -> torch.Tensor:
    if torch.is_tensor(x):
        return x.to(device="cuda", dtype=torch.float64).contiguous()
    return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64)).to("cuda")


def kappa_matrix(kappa_diff, kappa_idx, n: int) -> np.ndarray:
    """Antisymmetric kappa matrix of the non-redundant rotations (p < q)."""
    k = np.zeros((n, n))
    for d, (p, q) in zip(kappa_diff, kappa_idx):
        k[p, q] = d
        k[q, p] = -d
    return k


def rotate(c_old, kappa_mat: np.ndarray) -> torch.Tensor:
    """C_old expm(-kappa) (n_ao x n_mo) on the device."""
    require_gpu()
    c = _dev(c_old)
    n_ao, n = c.shape
    km = np.ascontiguousarray(kappa_mat, dtype=np.float64)
    if km.shape != (n, n):
        raise ValueError("kappa matrix has the wrong shape")
    out = torch.empty_like(c)
    N.check(N.lib().sq_orbital_rotation(n, n_ao, _ptr(c), km.ctypes.data, _ptr(out), _ptr(_work(n_ao, n)), _stream()))
    return out


def ao2mo(c, h_ao=None, g_ao=None) -> tuple[torch.Tensor | None, torch.Tensor | None]:
    """h_mo = C^T h C and g_mo[ijkl] = sum C_ai C_bj C_ck C_dl g_abcd on the device."""
    require_gpu()
    c = _dev(c)
    n_ao, n_mo = c.shape
    h = None if h_ao is None else _dev(h_ao)
    g = None if g_ao is None else _dev(g_ao)
    h_mo = None if h is None else torch.empty((n_mo, n_mo), dtype=torch.float64, device="cuda")
    g_mo = None if g is None else torch.empty((n_mo,) * 4, dtype=torch.float64, device="cuda")
    N.check(N.lib().sq_ao2mo(n_ao, n_mo, _ptr(c), _ptr(h), _ptr(g), _ptr(h_mo), _ptr(g_mo),
                             _ptr(_work(n_ao, n_mo)), _stream()))
    return h_mo, g_mo


def orbital_gradient(h_mo, g_mo, kappa_idx, num_inactive_orbs: int, num_active_orbs: int, rdm1, rdm2) -> np.ndarray:
    """<0|[kappa_mn, H]|0> for every (m, n) in kappa_idx (density_matrix.py:181-230)."""
    require_gpu()
    idx = np.ascontiguousarray(np.asarray(kappa_idx, dtype=np.int32).reshape(-1, 2))
    nk = len(idx)
    if nk == 0:
        return np.zeros(0)
    h = _dev(h_mo)
    g = _dev(g_mo)
    n = h.shape[0]
    r1 = _dev(rdm1)
    r2 = _dev(rdm2)
    grad = np.zeros(nk)
    N.check(N.lib().sq_orbital_gradient(n, num_inactive_orbs, num_active_orbs, _ptr(h), _ptr(g), _ptr(r1), _ptr(r2),
                                        nk, idx.ctypes.data, grad.ctypes.data,
                                        _ptr(_work(n, n, num_inactive_orbs, num_active_orbs, nk)), _stream()))
    return grad

