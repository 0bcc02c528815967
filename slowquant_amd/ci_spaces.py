"""CI_Info / get_indexing with the reference's interface (ci_spaces.py:9-116).

idx2det holds the reference's determinant integers (spin orbital k = 2p / 2p+1
at bit 2n-1-k; alpha strings outer, beta inner, lexicographic).  det2idx is a
read-only mapping computed on demand from string ranks instead of a hash map
of every determinant, so building it costs nothing even at 11.8 M
determinants.  The device engine builds its own tables (Space) with the same
order.
"""

from __future__ import annotations

import itertools
from collections.abc import Iterator, Mapping
from math import comb

import numpy as np


def _strings(n: int, k: int) -> np.ndarray:
    """Bit masks (bit p = orbital p) in itertools.combinations order."""
    if k < 0:
        return np.zeros(0, dtype=np.int64)
    return np.array([sum(1 << p for p in c) for c in itertools.combinations(range(n), k)], dtype=np.int64)


def _spread(masks: np.ndarray, n: int, beta: bool) -> np.ndarray:
    """Place spatial-orbital bits at interleaved positions 2n-1-(2p+beta)."""
    out = np.zeros_like(masks)
    for p in range(n):
        out |= ((masks >> p) & 1) << (2 * n - 1 - (2 * p + int(beta)))
    return out


def _rank(mask: int, n: int, k: int) -> int:
    """Lexicographic rank of a combination given as a bit mask."""
    r, prev, i = 0, -1, 0
    for c in range(n):
        if (mask >> c) & 1:
            for j in range(prev + 1, c):
                r += comb(n - 1 - j, k - 1 - i)
            prev = c
            i += 1
    return r


class _Det2Idx(Mapping):
    """det -> index for the fixed-(n, N_a, N_b) space, by closed-form ranking."""

    def __init__(self, n: int, na: int, nb: int) -> None:
        self.n, self.na, self.nb = n, na, nb
        self.nsb = comb(n, nb)
        self._len = comb(n, na) * self.nsb

    def _split(self, det: int):
        a = b = 0
        n = self.n
        for p in range(n):
            if (det >> (2 * n - 1 - 2 * p)) & 1:
                a |= 1 << p
            if (det >> (2 * n - 1 - (2 * p + 1))) & 1:
                b |= 1 << p
        return a, b

    def __getitem__(self, det) -> int:
        det = int(det)
        if det < 0 or det >> (2 * self.n):
            raise KeyError(det)
        a, b = self._split(det)
        if bin(a).count("1") != self.na or bin(b).count("1") != self.nb:
            raise KeyError(det)
        return _rank(a, self.n, self.na) * self.nsb + _rank(b, self.n, self.nb)

    def __contains__(self, det) -> bool:
        try:
            self[det]
            return True
        except KeyError:
            return False

    def __len__(self) -> int:
        return self._len

    def __iter__(self) -> Iterator[int]:
        a = _spread(_strings(self.n, self.na), self.n, False)
        b = _spread(_strings(self.n, self.nb), self.n, True)
        for x in a:
            for y in b:
                yield int(x | y)


class CI_Info:
    __slots__ = (
        "det2idx",
        "idx2det",
        "num_active_elec_alpha",
        "num_active_elec_beta",
        "num_active_orbs",
        "num_inactive_orbs",
        "num_virtual_orbs",
        "space_extension_offset",
    )

    def __init__(self, num_inactive_orbs, num_active_orbs, num_virtual_orbs, num_active_elec_alpha,
                 num_active_elec_beta, idx2det, det2idx) -> None:
        self.num_inactive_orbs = num_inactive_orbs
        self.num_active_orbs = num_active_orbs
        self.num_virtual_orbs = num_virtual_orbs
        self.num_active_elec_alpha = num_active_elec_alpha
        self.num_active_elec_beta = num_active_elec_beta
        self.idx2det = idx2det
        self.det2idx = det2idx
        self.space_extension_offset = 0


def get_indexing(num_inactive_orbs: int, num_active_orbs: int, num_virtual_orbs: int,
                 num_active_elec_alpha: int, num_active_elec_beta: int) -> CI_Info:
    n = num_active_orbs
    a = _spread(_strings(n, num_active_elec_alpha), n, False)
    b = _spread(_strings(n, num_active_elec_beta), n, True)
    idx2det = (a[:, None] | b[None, :]).reshape(-1).astype(np.int64)
    return CI_Info(num_inactive_orbs, num_active_orbs, num_virtual_orbs, num_active_elec_alpha,
                   num_active_elec_beta, idx2det, _Det2Idx(n, num_active_elec_alpha, num_active_elec_beta))
