"""Operator constructors (operators.py:8-529 of the reference): the excitation
generators of every gate type and the electronic Hamiltonian, as
FermionicOperator objects.  The GPU engine does not need these for its hot
path (it works from integral tensors and gate descriptors); they back the
generic ``propagate_state`` / ``expectation_value`` API.
"""

from __future__ import annotations

import numpy as np

from .fermionic_operator import FermionicOperator


def a_op(spinless_idx: int, spin: str, dagger: bool) -> FermionicOperator:
    if spin not in ("alpha", "beta"):
        raise ValueError(f'spin must be "alpha" or "beta" got {spin}')
    return FermionicOperator({((2 * spinless_idx + (spin == "beta"), dagger),): 1})


def a_op_spin(spin_idx: int, dagger: bool) -> FermionicOperator:
    return FermionicOperator({((spin_idx, dagger),): 1})


def Epq(p: int, q: int) -> FermionicOperator:
    """E_pq = a+_{p,a} a_{q,a} + a+_{p,b} a_{q,b}."""
    op = a_op(p, "alpha", True) * a_op(q, "alpha", False)
    op += a_op(p, "beta", True) * a_op(q, "beta", False)
    return op


def epqrs(p: int, q: int, r: int, s: int) -> FermionicOperator:
    """e_pqrs = E_pq E_rs - delta_qr E_ps."""
    op = Epq(p, q) * Epq(r, s)
    if q == r:
        op -= Epq(p, s)
    return op


def Eminuspq(p: int, q: int) -> FermionicOperator:
    return Epq(p, q) - Epq(q, p)


def commutator(A: FermionicOperator, B: FermionicOperator) -> FermionicOperator:
    return A * B - B * A


def double_commutator(A, B, C, do_symmetrized: bool = False) -> FermionicOperator:
    if do_symmetrized:
        return A * B * C + C * B * A - 1 / 2 * (A * C * B + B * C * A + C * A * B + B * A * C)
    return A * B * C - A * C * B - B * C * A + C * B * A


def _excitation(creators, annihilators, anti_hermitian: bool) -> FermionicOperator:
    op = a_op_spin(creators[0], True)
    for c in creators[1:]:
        op *= a_op_spin(c, True)
    for a in annihilators:
        op *= a_op_spin(a, False)
    if anti_hermitian:
        op -= op.dagger
    return op


def G1(i, a, return_anti_hermitian=False):
    """a+_a a_i (spin orbitals)."""
    return _excitation([a], [i], return_anti_hermitian)


def G2(i, j, a, b, return_anti_hermitian=False):
    """a+_a a+_b a_j a_i."""
    return _excitation([a, b], [j, i], return_anti_hermitian)


def G3(i, j, k, a, b, c, return_anti_hermitian=False):
    return _excitation([a, b, c], [k, j, i], return_anti_hermitian)


def G4(i, j, k, l, a, b, c, d, return_anti_hermitian=False):
    return _excitation([a, b, c, d], [l, k, j, i], return_anti_hermitian)


def G5(i, j, k, l, m, a, b, c, d, e, return_anti_hermitian=False):
    return _excitation([a, b, c, d, e], [m, l, k, j, i], return_anti_hermitian)


def G6(i, j, k, l, m, n, a, b, c, d, e, f, return_anti_hermitian=False):
    return _excitation([a, b, c, d, e, f], [n, m, l, k, j, i], return_anti_hermitian)


def G1_sa(i, a, return_anti_hermitian=False):
    """2^-1/2 E_ai."""
    op = 2 ** (-1 / 2) * Epq(a, i)
    if return_anti_hermitian:
        op -= op.dagger
    return op


def G2_sa(i, j, a, b, case, return_anti_hermitian=False):
    """Singlet spin-adapted doubles, cases 1-4 (symmetric) and 5 (antisymmetric)."""
    if case in (1, 2, 3, 4):
        fac = (2 if a == b else 1) * (2 if i == j else 1)
        op = 1 / 2 * fac ** (-1 / 2) * (Epq(a, i) * Epq(b, j) + Epq(a, j) * Epq(b, i))
    elif case == 5:
        op = 1 / (2 * 3 ** (1 / 2)) * (Epq(a, i) * Epq(b, j) - Epq(a, j) * Epq(b, i))
    else:
        raise ValueError(f"Got unknown case for spin-adapted doubles, {case}")
    if return_anti_hermitian:
        op -= op.dagger
    return op


def hamiltonian_full_space(h_mo: np.ndarray, g_mo: np.ndarray, num_orbs: int) -> FermionicOperator:
    H = FermionicOperator({})
    for p in range(num_orbs):
        for q in range(num_orbs):
            if abs(h_mo[p, q]) >= 1e-14:
                H += h_mo[p, q] * Epq(p, q)
    for p in range(num_orbs):
        for q in range(num_orbs):
            for r in range(num_orbs):
                for s in range(num_orbs):
                    if abs(g_mo[p, q, r, s]) >= 1e-14:
                        H += 1 / 2 * g_mo[p, q, r, s] * epqrs(p, q, r, s)
    return H


def hamiltonian_0i_0a(h_mo, g_mo, num_inactive_orbs: int, num_active_orbs: int) -> FermionicOperator:
    """Energy Hamiltonian over inactive + active orbitals (operators.py:476-529)."""
    H = FermionicOperator({})
    ni, na = num_inactive_orbs, num_active_orbs
    act = range(ni, ni + na)
    for i in range(ni):
        if abs(h_mo[i, i]) > 1e-14:
            H += h_mo[i, i] * Epq(i, i)
    for p in act:
        for q in act:
            if abs(h_mo[p, q]) > 1e-14:
                H += h_mo[p, q] * Epq(p, q)
    for i in range(ni):
        for j in range(ni):
            if abs(g_mo[i, i, j, j]) > 1e-14:
                H += 1 / 2 * g_mo[i, i, j, j] * epqrs(i, i, j, j)
            if i != j and abs(g_mo[j, i, i, j]) > 1e-14:
                H += 1 / 2 * g_mo[j, i, i, j] * epqrs(j, i, i, j)
    for i in range(ni):
        for p in act:
            for q in act:
                for idx in ((i, i, p, q), (p, q, i, i), (p, i, i, q), (i, p, q, i)):
                    if abs(g_mo[idx]) > 1e-14:
                        H += 1 / 2 * g_mo[idx] * epqrs(*idx)
    for p in act:
        for q in act:
            for r in act:
                for s in act:
                    if abs(g_mo[p, q, r, s]) > 1e-14:
                        H += 1 / 2 * g_mo[p, q, r, s] * epqrs(p, q, r, s)
    return H


def folded_integrals(h_mo: np.ndarray, g_mo: np.ndarray, num_inactive_orbs: int, num_active_orbs: int):
    """(e_core, h_eff, g_act) of hamiltonian_0i_0a folded to the active space.

    Doubly occupied inactive orbitals i contribute
      e_core = sum_i 2 h_ii + sum_ij (2 g_iijj - g_ijji)
      h_eff_pq = h_pq + sum_i (2 g_iipq - g_ipqi)
    Entries with |x| <= 1e-14 are dropped first, as hamiltonian_0i_0a does.
    """
    ni, na = num_inactive_orbs, num_active_orbs
    h = np.where(np.abs(h_mo) > 1e-14, h_mo, 0.0)
    g = np.where(np.abs(g_mo) > 1e-14, g_mo, 0.0)
    I = slice(0, ni)
    A = slice(ni, ni + na)
    e_core = 0.0
    if ni:
        gi = g[I, I, I, I]
        e_core = 2.0 * np.trace(h[I, I]) + 2.0 * np.einsum("iijj->", gi) - np.einsum("ijji->", gi)
    h_eff = np.array(h[A, A], dtype=np.float64)
    if ni:
        h_eff += 2.0 * np.einsum("iipq->pq", g[I, I, A, A]) - np.einsum("ipqi->pq", g[I, A, A, I])
    g_act = np.array(g[A, A, A, A], dtype=np.float64)
    return float(e_core), h_eff, g_act
