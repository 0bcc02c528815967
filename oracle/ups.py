"""UPS gate application and gradients, restated for the oracle (test only).

construct_ups_state      operator_state_algebra.py:963-1412
propagate_unitary        operator_state_algebra.py:1867-2309
get_grad_action          operator_state_algebra.py:2757-2865
gradient_reference       ups_wavefunction.py:1090-1141 (product-rule sweep)
gradient_adjoint         one forward sweep, one H.psi, one reverse sweep
gradient_plain           per-parameter derivative propagation (O(N^2))
"""

from __future__ import annotations

import numpy as np

from .apply import propagate_state
from .fermion import G1, G2, G3, G4, G5, G6, G2_sa, sa_double_tables

SINGLE_FREQ = ("single", "double", "triple", "quadruple", "quintuple", "sextuple", "sa_double_1")


def generator(exc_type: str, idx):
    """Anti-Hermitian generator T of one gate (operator_state_algebra.py:1046-1066)."""
    if exc_type == "single":
        return G1(*idx, ah=True)
    if exc_type == "double":
        return G2(*idx, ah=True)
    if exc_type == "triple":
        return G3(*idx, ah=True)
    if exc_type == "quadruple":
        return G4(*idx, ah=True)
    if exc_type == "quintuple":
        return G5(*idx, ah=True)
    if exc_type == "sextuple":
        return G6(*idx, ah=True)
    if exc_type.startswith("sa_double_"):
        case = int(exc_type[-1])
        return G2_sa(*idx, case, ah=True)
    raise ValueError(f"Got unknown excitation type, {exc_type}")


def _T(op, v, ci):
    return propagate_state([op], v, ci, do_folding=False)


def apply_gate(out, exc_type, idx, theta, ci):
    """exp(theta T) out by the reference's closed forms."""
    if abs(theta) < 1e-28:
        return out.copy()
    if exc_type == "sa_single":
        i, a = idx
        for Tg in (G1(2 * i, 2 * a, ah=True), G1(2 * i + 1, 2 * a + 1, ah=True)):
            t1 = _T(Tg, out, ci)
            t2 = _T(Tg, t1, ci)
            out = out + np.sin(theta) * t1 + (1 - np.cos(theta)) * t2
        return out
    if exc_type in SINGLE_FREQ:
        Tg = generator(exc_type, idx)
        t1 = _T(Tg, out, ci)
        t2 = _T(Tg, t1, ci)
        return out + np.sin(theta) * t1 + (1 - np.cos(theta)) * t2
    if exc_type in ("sa_double_2", "sa_double_3", "sa_double_4", "sa_double_5"):
        Tg = generator(exc_type, idx)
        S, K = sa_double_tables()[int(exc_type[-1])]
        out = out.copy()
        tmp = out
        for k, row in enumerate(K, start=1):
            tmp = _T(Tg, tmp, ci)
            if k % 2 == 1:
                c = sum(kf * np.sin(sf * theta) for kf, sf in zip(row, S))
            else:
                c = sum(kf * (np.cos(sf * theta) - 1) for kf, sf in zip(row, S))
            out = out + c * tmp
        return out
    raise ValueError(f"Got unknown excitation type, {exc_type}")


def construct_ups_state(state, ci, thetas, layout, dagger=False):
    out = np.array(state, dtype=np.float64, copy=True)
    order = -1 if dagger else 1
    for t, idx, th in zip(layout.excitation_operator_type[::order], layout.excitation_indices[::order], list(thetas)[::order]):
        if abs(th) < 1e-28:
            continue
        out = apply_gate(out, t, idx, -th if dagger else th, ci)
    return out


def propagate_unitary(state, i, ci, thetas, layout):
    return apply_gate(np.asarray(state, dtype=np.float64), layout.excitation_operator_type[i], layout.excitation_indices[i], thetas[i], ci)


def get_grad_action(state, i, ci, layout):
    t = layout.excitation_operator_type[i]
    idx = layout.excitation_indices[i]
    if t == "sa_single":
        a, b = idx
        T = G1(2 * a, 2 * b, ah=True) + G1(2 * a + 1, 2 * b + 1, ah=True)
        return _T(T, state, ci)
    return _T(generator(t, idx), state, ci)


def gradient_reference(H, csf, ci, thetas, layout):
    """Product-rule sweep of ups_wavefunction.py:1090-1141 (H already folded or active)."""
    psi = construct_ups_state(csf, ci, thetas, layout)
    bra = propagate_state([H], psi, ci, do_folding=False)
    bra = construct_ups_state(bra, ci, thetas, layout, dagger=True)
    ket = csf.copy()
    g = np.zeros(len(thetas))
    for i in range(len(thetas)):
        g[i] = 2 * bra @ get_grad_action(ket, i, ci, layout)
        bra = propagate_unitary(bra, i, ci, thetas, layout)
        ket = propagate_unitary(ket, i, ci, thetas, layout)
    return g


def gradient_adjoint(H, csf, ci, thetas, layout):
    """Reverse sweep: g_i = 2 <lambda_i|T_i psi_i>, un-applying gate i on both."""
    psi = construct_ups_state(csf, ci, thetas, layout)
    lam = propagate_state([H], psi, ci, do_folding=False)
    E = float(psi @ lam)
    g = np.zeros(len(thetas))
    neg = [-t for t in thetas]
    for i in range(len(thetas) - 1, -1, -1):
        g[i] = 2 * lam @ get_grad_action(psi, i, ci, layout)
        psi = propagate_unitary(psi, i, ci, neg, layout)
        lam = propagate_unitary(lam, i, ci, neg, layout)
    return E, g


def gradient_plain(H, csf, ci, thetas, layout):
    """g_i = 2 <H psi | U_N..U_{i+1} U_i T_i U_{i-1}..U_1 |csf>, each from scratch."""
    psi = construct_ups_state(csf, ci, thetas, layout)
    lam = propagate_state([H], psi, ci, do_folding=False)
    g = np.zeros(len(thetas))
    N = len(thetas)
    for i in range(N):
        v = csf.copy()
        for k in range(N):
            v = propagate_unitary(v, k, ci, thetas, layout)
            if k == i:
                v = get_grad_action(v, i, ci, layout)
        g[i] = 2 * lam @ v
    return g
