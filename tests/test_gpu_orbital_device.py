"""Orbital rotation, AO -> MO transforms and the orbital gradient on the
device (sq_orbital_rotation, sq_ao2mo, sq_orbital_gradient) against the
host restatements:

* c_mo = C expm(-kappa) (ups_wavefunction.py:367-385) against
  scipy.linalg.expm, for small and large rotations (the latter exercise the
  squarings) and a rectangular C (more AOs than MOs), on both sides of the
  one-block limit (48 orbitals);
* h_mo, g_mo (integralfunctions.py:186-211) against numpy.einsum, with
  n_ao = n_mo and n_ao > n_mo;
* get_orbital_gradient (density_matrix.py:181-230) against the oracle's loop
  restatement (oracle/orbital.py) for random Gamma-shaped inputs over
  inactive, active and virtual orbitals.

Tolerance 1e-12 relative to the largest element.
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.linalg

pytestmark = pytest.mark.gpu


def _sym_g(n, rng):
    B = rng.normal(size=(3, n, n))
    B = B + B.transpose(0, 2, 1)
    return np.einsum("Lpq,Lrs->pqrs", B, B)


# n <= 48 runs in one block (k_expm_small), larger n through GEMM launches
@pytest.mark.parametrize("n,n_ao,scale", [(2, 3, 0.3), (6, 6, 0.05), (10, 10, 3.0), (7, 9, 0.4), (33, 33, 1.0), (48, 50, 1.0),
                                          (49, 49, 0.5), (64, 70, 2.0)])
def test_orbital_rotation_against_expm(n, n_ao, scale):
    from slowquant_amd import orbitals

    rng = np.random.default_rng(n + n_ao)
    c_old = rng.normal(size=(n_ao, n))
    idx = [(p, q) for p in range(n) for q in range(p + 1, n)]
    diff = rng.normal(0.0, scale, len(idx))
    km = orbitals.kappa_matrix(diff, idx, n)
    got = orbitals.rotate(c_old, km).cpu().numpy()
    expect = c_old @ scipy.linalg.expm(-km)
    assert np.abs(got - expect).max() < 1e-12 * max(1.0, np.abs(expect).max())


# n_ao, n_mo <= 8 run in one block (k_ao2mo_small), larger through GEMM launches
@pytest.mark.parametrize("n_ao,n_mo", [(2, 1), (4, 4), (8, 8), (8, 6), (9, 9), (10, 10), (12, 9), (20, 20)])
def test_ao2mo_against_einsum(n_ao, n_mo):
    from slowquant_amd import orbitals

    rng = np.random.default_rng(7 * n_ao + n_mo)
    c = rng.normal(size=(n_ao, n_mo))
    h = rng.normal(size=(n_ao, n_ao))
    h = h + h.T
    g = _sym_g(n_ao, rng)
    h_mo, g_mo = orbitals.ao2mo(c, h, g)
    eh = np.einsum("ai,bj,ab->ij", c, c, h, optimize=True)
    eg = np.einsum("ai,bj,ck,dl,abcd->ijkl", c, c, c, c, g, optimize=True)
    assert np.abs(h_mo.cpu().numpy() - eh).max() < 1e-12 * np.abs(eh).max()
    assert np.abs(g_mo.cpu().numpy() - eg).max() < 1e-12 * np.abs(eg).max()
    # either transform alone
    h2, g2 = orbitals.ao2mo(c, h, None)
    assert g2 is None and np.abs(h2.cpu().numpy() - eh).max() < 1e-12 * np.abs(eh).max()


# inactive + active <= 8 runs in one block (k_kappa_grad_small), more through GEMM launches
@pytest.mark.parametrize("ni,na,nv", [(0, 1, 1), (0, 4, 0), (2, 2, 2), (1, 4, 1), (3, 5, 4), (2, 7, 3), (4, 6, 6)])
def test_orbital_gradient_against_oracle(ni, na, nv):
    from oracle.orbital import get_orbital_gradient as oracle_gradient
    from slowquant_amd import orbitals

    n = ni + na + nv
    rng = np.random.default_rng(100 + n)
    h = rng.normal(size=(n, n))
    h = h + h.T
    g = _sym_g(n, rng)
    r1 = rng.normal(size=(na, na))
    r1 = r1 + r1.T
    r2 = rng.normal(size=(na,) * 4)
    idx = [(p, q) for p in range(n) for q in range(p + 1, n)]
    got = orbitals.orbital_gradient(h, g, idx, ni, na, r1, r2)
    expect = oracle_gradient(h, g, idx, ni, na, r1, r2)
    assert np.abs(got - expect).max() < 1e-12 * max(1.0, np.abs(expect).max())
