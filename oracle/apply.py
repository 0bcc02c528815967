"""Determinant scans and propagate_state, restated for the oracle (test only).

``apply_string`` restates the gather form of the reference's Numba kernel
apply_operator_threaded (operator_state_algebra.py:139-219): for every target
determinant that survives the creator / screened-annihilator masks, the source
determinant is found by flipping the creator bits then the annihilator bits,
the Jordan-Wigner phase is accumulated with parity_check, and
``out[i] += sign * factor * state[source]``.  numpy vectorises the loop over
determinants; the per-string structure is the reference's.
"""

from __future__ import annotations

import numpy as np

from .ci import CIInfo, parity_check
from .fermion import FOp


def _popcount(x: np.ndarray) -> np.ndarray:
    return np.bitwise_count(x.astype(np.uint64)).astype(np.int64)


def string_masks(key, n_act):
    """Split a normal-ordered string as propagate_state does (operator_state_algebra.py:592-612)."""
    nb1 = 2 * n_act - 1
    create = [i for i, d in key if d]
    anni = [i for i, d in key if not d]
    anni_screen = [i for i in anni if i not in create]
    cmask = 0
    for i in create:
        cmask |= 1 << (nb1 - i)
    amask = 0
    for i in anni_screen:
        amask |= 1 << (nb1 - i)
    return create + anni, cmask, amask


def apply_string(state, ci: CIInfo, key, factor, out, pc=None, do_unsafe=False):
    n_act = ci.num_active_orbs
    nb1 = 2 * n_act - 1
    if pc is None:
        pc = parity_check(n_act)
    a_string, cmask, amask = string_masks(key, n_act)
    det = ci.idx2det
    ok = ((det & cmask) == cmask) & ((det & amask) == 0)
    rows = np.nonzero(ok)[0]
    d = det[rows].copy()
    phase = np.zeros(len(d), dtype=np.int64)
    for orb in a_string:
        d ^= np.int64(1) << np.int64(nb1 - orb)
        phase += _popcount(d & pc[orb])
    src = ci.lookup(d)
    if (src < 0).any():
        if not do_unsafe:
            raise KeyError("operator maps outside the CI space")
        keep = src >= 0
        rows, src, phase = rows[keep], src[keep], phase[keep]
    sign = 1.0 - 2.0 * (phase & 1)
    np.add.at(out, rows, sign * factor * state[src])
    return out


def propagate_state(operators, state, ci: CIInfo, do_folding=True, do_unsafe=False):
    """FermionicOperator part of propagate_state (operator_state_algebra.py:472-630).

    Operators are applied right to left; each may be folded to the active space.
    """
    new = np.array(state, dtype=np.float64, copy=True)
    pc = parity_check(ci.num_active_orbs)
    for op in operators[::-1]:
        if do_folding:
            op = op.folded(ci.num_inactive_orbs, ci.num_active_orbs, ci.num_virtual_orbs)
        tmp = np.zeros_like(new)
        for key, f in op.operators.items():
            apply_string(new, ci, key, f, tmp, pc, do_unsafe)
        new = tmp
    return new


def expectation_value(bra, operators, ket, ci: CIInfo, do_folding=True):
    """bra @ propagate_state(operators, ket) (operator_state_algebra.py:784-824)."""
    return float(bra @ propagate_state(operators, ket, ci, do_folding))


def operator_matrix(op: FOp, ci: CIInfo) -> np.ndarray:
    """Dense matrix of an active-space operator via the same scans (operator_state_algebra.py:413-469)."""
    d = len(ci.idx2det)
    M = np.zeros((d, d))
    eye = np.eye(d)
    for j in range(d):
        M[:, j] = propagate_state([op], eye[:, j], ci, do_folding=False)
    return M
