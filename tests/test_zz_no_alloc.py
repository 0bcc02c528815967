"""sq_sigma_noalloc (sigma = H psi with its scratch in a file-scope device
array, for callers that cannot allocate) against sq_sigma with a caller
workspace, through the C ABI: two calls in a row, the second on a smaller
space, then the first space again (the shared scratch keeps the previous
call's contents).  Tolerance: 1e-12 relative to max|sigma| element-wise,
energies 1e-10.  A space whose scratch exceeds the 64 MiB array is refused
with ValueError before any launch.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(sp, H, psi):
    import torch

    from slowquant_amd import _native as N

    L = N.lib()
    fn = L.sq_sigma_noalloc
    fn.argtypes = [ctypes.c_void_p] * 6
    fn.restype = ctypes.c_int
    out = torch.empty_like(psi)
    e = torch.zeros(1, dtype=torch.float64, device="cuda")
    N.check(fn(sp.handle, H.handle, psi.data_ptr(), out.data_ptr(), e.data_ptr(),
               torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return out, float(e.item())


def _case(n, na, nb, seed):
    import torch

    from slowquant_amd import engine as E
    from slowquant_amd.synthetic import synthetic_problem

    h, g, _ = synthetic_problem(n, 2 * n, 1, 0.05, seed)
    sp = E.Space(n, na, nb)
    H = E.Hamiltonian(sp, h, g, e_core=0.25)
    rng = np.random.default_rng(seed)
    psi = torch.tensor(rng.normal(size=sp.dim), dtype=torch.float64, device="cuda")
    psi /= torch.linalg.norm(psi)
    return sp, H, psi


def _check(sp, H, psi):
    ref, e_ref = H.sigma(psi)
    got, e = _noalloc(sp, H, psi)
    scale = max(1.0, float(ref.abs().max()))
    print(f"noalloc sigma dim {sp.dim}: max diff {float((got - ref).abs().max()):.2e}, dE {abs(e - e_ref):.2e}",
          flush=True)
    assert float((got - ref).abs().max()) <= 1e-12 * scale
    assert abs(e - e_ref) <= 1e-10


def test_sigma_noalloc_matches_sigma():
    big = _case(12, 6, 6, 5)
    small = _case(8, 4, 4, 6)
    _check(*big)
    _check(*small)  # second call, smaller input
    _check(*big)


def test_sigma_noalloc_rejects_oversized_space():
    from slowquant_amd import engine as E

    sp, H, psi = _case(6, 3, 3, 7)
    big = E.Space(14, 7, 7)  # 11,778,624 determinants: more scratch than the array holds
    h = np.eye(14)
    Hb = E.Hamiltonian(big, h, np.zeros((14,) * 4))
    import torch

    psi_b = torch.zeros(big.dim, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        _noalloc(big, Hb, psi_b)
    _check(sp, H, psi)  # still usable after the refusal
