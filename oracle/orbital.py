"""Orbital rotations, integral transforms and the orbital gradient, restated
for the oracle -- TEST INFRASTRUCTURE ONLY.

c_mo                      ups_wavefunction.py:367-385 (C expm(-kappa))
one/two_electron_integral_transform
                          integralfunctions.py:186-211 (einsum contractions)
RDM1 / RDM2               density_matrix.py:5-136 (full-space closed forms)
get_electronic_energy     density_matrix.py:139-178
get_orbital_gradient      density_matrix.py:181-230 (the reference's loops,
                          element by element)
energy_at                 E(kappa, theta) of a UPS or UCC state: folded
                          Hamiltonian (hamiltonian_0i_0a + folding) at the
                          rotated orbitals, the oracle's state, <psi|H|psi>

Pinned by: get_orbital_gradient against central differences of energy_at in
kappa (tests/test_oracle.py), get_electronic_energy against <psi|H|psi> of the
folded Hamiltonian, and the transforms against explicit loops on tiny bases.
"""

from __future__ import annotations

import numpy as np
import scipy.linalg

from . import fermion as F
from .apply import expectation_value
from .ci import get_indexing


def c_mo(c_ref: np.ndarray, kappa, kappa_idx) -> np.ndarray:
    """C expm(-K) with K[p,q] = kappa, K[q,p] = -kappa (ups_wavefunction.py:367-385)."""
    K = np.zeros_like(c_ref)
    for v, (p, q) in zip(kappa, kappa_idx):
        K[p, q] = v
        K[q, p] = -v
    return c_ref @ scipy.linalg.expm(-K)


def one_electron_integral_transform(C, h_ao):
    return np.einsum("ai,bj,ab->ij", C, C, h_ao)


def two_electron_integral_transform(C, g_ao):
    return np.einsum("ai,bj,ck,dl,abcd->ijkl", C, C, C, C, g_ao, optimize=True)


def RDM1(p, q, ni, na, rdm1):
    virt = ni + na
    if p >= virt or q >= virt:
        return 0.0
    if p >= ni and q >= ni:
        return rdm1[p - ni, q - ni]
    if p < ni and q < ni:
        return 2.0 if p == q else 0.0
    return 0.0


def RDM2(p, q, r, s, ni, na, rdm1, rdm2):
    virt = ni + na
    if p >= virt or q >= virt or r >= virt or s >= virt:
        return 0.0
    ap, aq, ar, as_ = p >= ni, q >= ni, r >= ni, s >= ni
    if ap and aq and ar and as_:
        return rdm2[p - ni, q - ni, r - ni, s - ni]
    if not ap and aq and ar and not as_:
        return -rdm1[q - ni, r - ni] if p == s else 0.0
    if ap and not aq and not ar and as_:
        return -rdm1[p - ni, s - ni] if q == r else 0.0
    if ap and aq and not ar and not as_:
        return 2.0 * rdm1[p - ni, q - ni] if r == s else 0.0
    if not ap and not aq and ar and as_:
        return 2.0 * rdm1[r - ni, s - ni] if p == q else 0.0
    if not (ap or aq or ar or as_):
        v = 0.0
        if p == q and r == s:
            v += 4.0
        if q == r and p == s:
            v -= 2.0
        return v
    return 0.0


def get_electronic_energy(h, g, ni, na, rdm1, rdm2) -> float:
    o = ni + na
    e = 0.0
    for p in range(o):
        for q in range(o):
            e += h[p, q] * RDM1(p, q, ni, na, rdm1)
    for p in range(o):
        for q in range(o):
            for r in range(o):
                for s in range(o):
                    e += 0.5 * g[p, q, r, s] * RDM2(p, q, r, s, ni, na, rdm1, rdm2)
    return e


def get_orbital_gradient(h, g, kappa_idx, ni, na, rdm1, rdm2) -> np.ndarray:
    """<0|[kappa_mn, H]|0> by the loops of density_matrix.py:209-229."""
    o = ni + na
    out = np.zeros(len(kappa_idx))
    for idx, (m, n) in enumerate(kappa_idx):
        v = 0.0
        for p in range(o):
            v += 2 * h[n, p] * RDM1(m, p, ni, na, rdm1)
            v -= 2 * h[p, m] * RDM1(p, n, ni, na, rdm1)
        for p in range(o):
            for q in range(o):
                for r in range(o):
                    v += g[n, p, q, r] * RDM2(m, p, q, r, ni, na, rdm1, rdm2)
                    v -= g[p, m, q, r] * RDM2(p, n, q, r, ni, na, rdm1, rdm2)
                    v -= g[m, p, q, r] * RDM2(n, p, q, r, ni, na, rdm1, rdm2)
                    v += g[p, n, q, r] * RDM2(p, m, q, r, ni, na, rdm1, rdm2)
        out[idx] = v
    return out


def folded_hamiltonian(h_mo, g_mo, ni, na, nv):
    """hamiltonian_0i_0a folded to the active space (operators.py:476-529,
    fermionic_operator.py:379-471)."""
    H = F.hamiltonian_0i_0a(h_mo, g_mo, ni, na)
    return H.folded(ni, na, nv)


def energy_at(c_ref, h_ao, g_ao, kappa, kappa_idx, state_fn, ni, na, nv, nalpha, nbeta):
    """E(kappa) = <psi|H(kappa)|psi>, H folded at the orbitals C expm(-kappa);
    state_fn(ci) returns the active-space state (reference convention)."""
    C = c_mo(c_ref, kappa, kappa_idx)
    h = one_electron_integral_transform(C, h_ao)
    g = two_electron_integral_transform(C, g_ao)
    H = folded_hamiltonian(h, g, ni, na, nv)
    ci = get_indexing(ni, na, nv, nalpha, nbeta)
    psi = state_fn(ci)
    return expectation_value(psi, [H], psi, ci, do_folding=False), h, g, psi, ci
