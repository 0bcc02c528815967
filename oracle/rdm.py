"""Active-space 1-/2-RDMs by expectation values, restated for the oracle (test only).

Loop structure and symmetry fills follow ups_wavefunction.py:409-476
(identical in ucc_wavefunction.py:339-406, which applies the two E_pq in turn).
Active indices are used directly (no inactive offset) on an active-space CI.
"""

from __future__ import annotations

import numpy as np

from .apply import propagate_state
from .fermion import Epq


def rdm1(psi, ci):
    n = ci.num_active_orbs
    out = np.zeros((n, n))
    for p in range(n):
        for q in range(p + 1):
            v = float(psi @ propagate_state([Epq(p, q)], psi, ci, do_folding=False))
            out[p, q] = v
            out[q, p] = v
    return out


def rdm2_loop_indices(n):
    """The (p, q, r, s) visited by ups_wavefunction.py:444-469."""
    for p in range(n):
        for q in range(p + 1):
            for r in range(p + 1):
                if p == q:
                    s_lim = r + 1
                elif p == r:
                    s_lim = q + 1
                elif q < r:
                    s_lim = p
                else:
                    s_lim = p + 1
                for s in range(s_lim):
                    yield p, q, r, s


def rdm2(psi, ci, r1=None):
    n = ci.num_active_orbs
    if r1 is None:
        r1 = rdm1(psi, ci)
    out = np.zeros((n, n, n, n))
    for p, q, r, s in rdm2_loop_indices(n):
        v = float(psi @ propagate_state([Epq(p, q), Epq(r, s)], psi, ci, do_folding=False))
        if q == r:
            v -= r1[p, s]
        out[p, q, r, s] = v
        out[r, s, p, q] = v
        out[q, p, s, r] = v
        out[s, r, q, p] = v
    return out
