"""Brute-force second quantisation on the CI space (test infrastructure only).

Independent of ``oracle.fermion`` / ``oracle.apply``: determinants are sets of
occupied spin orbitals, operators are applied literally right to left with the
textbook Jordan-Wigner phase (-1)^(#occupied spin orbitals with lower index),
with no normal ordering.  Used to pin the scan kernels, the normal-ordering
code and every gate generator.
"""

from __future__ import annotations

import itertools

import numpy as np


def ci_dets(n_act: int, na: int, nb: int) -> list[frozenset]:
    """Determinants in the reference's index order (alpha outer, beta inner)."""
    out = []
    for a in itertools.combinations(range(n_act), na):
        for b in itertools.combinations(range(n_act), nb):
            out.append(frozenset([2 * p for p in a] + [2 * p + 1 for p in b]))
    return out


def apply_product(ops, det: frozenset):
    """Apply ops = [(spin_orb, is_creator), ...] (leftmost applied last)."""
    occ = set(det)
    sign = 1
    for k, dag in reversed(ops):
        below = sum(1 for q in occ if q < k)
        if dag:
            if k in occ:
                return 0, None
            occ.add(k)
        else:
            if k not in occ:
                return 0, None
            occ.remove(k)
        if below % 2:
            sign = -sign
    return sign, frozenset(occ)


def matrix(terms, n_act: int, na: int, nb: int) -> np.ndarray:
    """Dense matrix of sum_c c * product(ops) for terms = [(c, ops), ...]."""
    dets = ci_dets(n_act, na, nb)
    index = {d: i for i, d in enumerate(dets)}
    M = np.zeros((len(dets), len(dets)))
    for j, d in enumerate(dets):
        for c, ops in terms:
            s, nd = apply_product(ops, d)
            if s == 0:
                continue
            M[index[nd], j] += c * s
    return M


def cre(k):
    return (k, True)


def ann(k):
    return (k, False)


def E_terms(p, q, c=1.0):
    return [(c, [cre(2 * p), ann(2 * q)]), (c, [cre(2 * p + 1), ann(2 * q + 1)])]


def product_terms(A, B):
    return [(ca * cb, oa + ob) for ca, oa in A for cb, ob in B]


def dagger_terms(A):
    return [(c, [(k, not d) for k, d in reversed(ops)]) for c, ops in A]


def anti_hermitian(A):
    return A + [(-c, ops) for c, ops in dagger_terms(A)]


def hamiltonian_terms(h, g, n):
    """H = sum h_pq E_pq + 1/2 sum g_pqrs (E_pq E_rs - delta_qr E_ps), literal products."""
    T = []
    for p in range(n):
        for q in range(n):
            T += E_terms(p, q, h[p, q])
    for p in range(n):
        for q in range(n):
            for r in range(n):
                for s in range(n):
                    c = 0.5 * g[p, q, r, s]
                    T += product_terms(E_terms(p, q, c), E_terms(r, s))
                    if q == r:
                        T += E_terms(p, s, -c)
    return T
