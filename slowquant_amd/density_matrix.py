"""Full-space density matrices, RDM energy and orbital gradient (host, numpy).

Same element rules as density_matrix.py:5-230 of the reference (RDM1, RDM2,
get_electronic_energy, get_orbital_gradient), evaluated with dense numpy
contractions over zero-padded full-space Gamma tensors instead of scalar
loops.  The active-space rdm1/rdm2 they consume come from the GPU engine.
"""

from __future__ import annotations

import numpy as np


def RDM1(p: int, q: int, num_inactive_orbs: int, num_active_orbs: int, rdm1: np.ndarray) -> float:
    ni, na = num_inactive_orbs, num_active_orbs
    if p >= ni + na or q >= ni + na:
        return 0
    if p >= ni and q >= ni:
        return rdm1[p - ni, q - ni]
    if p < ni and q < ni:
        return 2 if p == q else 0
    return 0


def RDM2(p, q, r, s, num_inactive_orbs, num_active_orbs, rdm1, rdm2) -> float:
    ni, na = num_inactive_orbs, num_active_orbs
    v = ni + na
    if p >= v or q >= v or r >= v or s >= v:
        return 0
    ap, aq, ar, as_ = p >= ni, q >= ni, r >= ni, s >= ni
    if ap and aq and ar and as_:
        return rdm2[p - ni, q - ni, r - ni, s - ni]
    if not ap and aq and ar and not as_:
        return -rdm1[q - ni, r - ni] if p == s else 0
    if ap and not aq and not ar and as_:
        return -rdm1[p - ni, s - ni] if q == r else 0
    if ap and aq and not ar and not as_:
        return 2 * rdm1[p - ni, q - ni] if r == s else 0
    if not ap and not aq and ar and as_:
        return 2 * rdm1[r - ni, s - ni] if p == q else 0
    if not (ap or aq or ar or as_):
        val = 0
        if p == q and r == s:
            val += 4
        if q == r and p == s:
            val -= 2
        return val
    return 0


def full_space_rdms(num_inactive_orbs: int, num_active_orbs: int, rdm1: np.ndarray, rdm2: np.ndarray, n_total=None):
    """Gamma1 (N x N) and Gamma2 (N^4) over N = n_total orbitals (default: inactive + active)."""
    ni, na = num_inactive_orbs, num_active_orbs
    o = ni + na
    N = o if n_total is None else n_total
    G1 = np.zeros((N, N))
    G2 = np.zeros((N, N, N, N))
    A = slice(ni, o)
    I = np.arange(ni)
    G1[I, I] = 2.0
    G1[A, A] = rdm1
    G2[A, A, A, A] = rdm2
    for i in range(ni):
        G2[i, A, A, i] = -rdm1
        G2[A, i, i, A] = -rdm1
        G2[A, A, i, i] = 2 * rdm1
        G2[i, i, A, A] = 2 * rdm1
        for j in range(ni):
            G2[i, i, j, j] += 4.0
            G2[i, j, j, i] -= 2.0
    return G1, G2


def get_electronic_energy(h_int, g_int, num_inactive_orbs, num_active_orbs, rdm1, rdm2) -> float:
    """E = sum h Gamma1 + 1/2 sum g Gamma2 over inactive + active orbitals."""
    o = num_inactive_orbs + num_active_orbs
    G1, G2 = full_space_rdms(num_inactive_orbs, num_active_orbs, rdm1, rdm2)
    return float(np.sum(h_int[:o, :o] * G1) + 0.5 * np.sum(g_int[:o, :o, :o, :o] * G2))


def get_orbital_gradient(h_int, g_int, kappa_idx, num_inactive_orbs, num_active_orbs, rdm1, rdm2) -> np.ndarray:
    """<0|[kappa_mn, H]|0> for every (m, n) in kappa_idx (density_matrix.py:181-230)."""
    kappa_idx = np.asarray(kappa_idx, dtype=int).reshape(-1, 2)
    if len(kappa_idx) == 0:
        return np.zeros(0)
    N = h_int.shape[0]
    o = num_inactive_orbs + num_active_orbs
    G1, G2 = full_space_rdms(num_inactive_orbs, num_active_orbs, rdm1, rdm2, n_total=N)
    # p, q, r run over the inactive + active orbitals only (Gamma is zero outside).
    h = h_int[:, :o]
    g = g_int[:, :o, :o, :o]
    G1c = G1[:, :o]
    G2c = G2[:, :o, :o, :o]
    Y = h @ G1c.T  # Y[n, m] = sum_p h[n,p] G1[m,p]
    V = h_int[:o, :].T @ G1[:o, :]  # V[m, n] = sum_p h[p,m] G1[p,n]
    Z = np.einsum("apqr,bpqr->ab", g, G2c, optimize=True)  # Z[a,b] = sum g[a,pqr] G2[b,pqr]
    gT = g_int[:o, :, :o, :o]
    G2T = G2[:o, :, :o, :o]
    W = np.einsum("paqr,pbqr->ab", gT, G2T, optimize=True)  # W[a,b] = sum g[p,a,q,r] G2[p,b,q,r]
    m = kappa_idx[:, 0]
    n = kappa_idx[:, 1]
    return 2 * Y[n, m] - 2 * V[m, n] + Z[n, m] - W[m, n] - Z[m, n] + W[n, m]
