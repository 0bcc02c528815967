"""Determinant addressing, restated from ci_spaces.py:56-116 (oracle only).

Determinant integer convention (ci_spaces.py:101-104): 2n bits, spin orbital
k = 2p (alpha) / 2p+1 (beta) sits at bit (2n-1-k).  Index order: alpha strings
outer, beta strings inner, each in itertools.combinations lexicographic order.
"""

from __future__ import annotations

import itertools

import numpy as np


def spin_strings(num_orbs: int, num_elec: int) -> list[tuple[int, ...]]:
    """Occupied-orbital tuples in itertools.combinations order (ci_spaces.py:56-73)."""
    if num_elec < 0:
        return []
    return list(itertools.combinations(range(num_orbs), num_elec))


class CIInfo:
    """Plain container mirroring CI_Info (ci_spaces.py:9-53)."""

    def __init__(self, n_inact, n_act, n_virt, na, nb, idx2det, det2idx):
        self.num_inactive_orbs = n_inact
        self.num_active_orbs = n_act
        self.num_virtual_orbs = n_virt
        self.num_active_elec_alpha = na
        self.num_active_elec_beta = nb
        self.idx2det = idx2det
        self.det2idx = det2idx
        self.space_extension_offset = 0
        # Vectorised lookup: idx2det is not sorted, so keep a sorted copy.
        order = np.argsort(idx2det)
        self._sorted_dets = idx2det[order]
        self._sorted_idx = order

    def lookup(self, dets: np.ndarray) -> np.ndarray:
        """det -> index for an array of dets that are all in the space."""
        pos = np.searchsorted(self._sorted_dets, dets)
        pos = np.minimum(pos, len(self._sorted_dets) - 1)
        ok = self._sorted_dets[pos] == dets
        out = np.where(ok, self._sorted_idx[pos], -1)
        return out


def get_indexing(n_inact: int, n_act: int, n_virt: int, na: int, nb: int) -> CIInfo:
    """Restates get_indexing (ci_spaces.py:76-116) with integer arithmetic.

    The reference builds the interleaved bit string as text; here the same
    integer is assembled from bit positions.
    """
    nbits = 2 * n_act
    idx2det = []
    det2idx = {}
    idx = 0
    for a in spin_strings(n_act, na):
        amask = 0
        for p in a:
            amask |= 1 << (nbits - 1 - 2 * p)
        for b in spin_strings(n_act, nb):
            det = amask
            for p in b:
                det |= 1 << (nbits - 1 - (2 * p + 1))
            idx2det.append(det)
            det2idx[det] = idx
            idx += 1
    return CIInfo(n_inact, n_act, n_virt, na, nb, np.array(idx2det, dtype=np.int64), det2idx)


def parity_check(n_act: int) -> np.ndarray:
    """parity_check[k] masks spin orbitals 0..k-1 (operator_state_algebra.py:517-522)."""
    pc = np.zeros(2 * n_act + 1, dtype=np.int64)
    num = 0
    for i in range(2 * n_act - 1, -1, -1):
        num += 2**i
        pc[2 * n_act - i] = num
    return pc
