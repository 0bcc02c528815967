"""Unfactorised UCC state, restated for the oracle (test only).

get_ucc_T / construct_ucc_state: operator_state_algebra.py:870-960, with the
dense matrix of build_operator_matrix (:413-469) and
scipy.sparse.linalg.expm_multiply(T, psi, traceA=0) (:893-896).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse.linalg as ssl

from .apply import operator_matrix
from .fermion import G1, G2, G3, G4, G5, G6, G1_sa, G2_sa, FOp


def get_ucc_T(thetas, layout) -> FOp:
    T = FOp({})
    for t, idx, th in zip(layout.excitation_operator_type, layout.excitation_indices, thetas):
        if abs(th) < 1e-28:
            continue
        if t == "sa_single":
            op = G1_sa(*idx, ah=True)
        elif t.startswith("sa_double_"):
            op = G2_sa(*idx, int(t[-1]), ah=True)
        elif t == "single":
            op = G1(*idx, ah=True)
        elif t == "double":
            op = G2(*idx, ah=True)
        elif t == "triple":
            op = G3(*idx, ah=True)
        elif t == "quadruple":
            op = G4(*idx, ah=True)
        elif t == "quintuple":
            op = G5(*idx, ah=True)
        elif t == "sextuple":
            op = G6(*idx, ah=True)
        else:
            raise ValueError(t)
        T = T + float(th) * op
    return T


def construct_ucc_state(state, ci, thetas, layout, dagger=False):
    Tmat = operator_matrix(get_ucc_T(thetas, layout), ci)
    if dagger:
        Tmat = -Tmat
    return ssl.expm_multiply(Tmat, np.asarray(state, dtype=np.float64), traceA=0.0)


def construct_ucc_state_matfree(state, ci, thetas, layout, nthreads=1):
    """construct_ucc_state without the dense matrix: the same generator
    get_ucc_T applied by the C scan (oracle/csrc/scan.c, the restatement of
    operator_state_algebra.py:139-219) inside a LinearOperator, exponentiated
    by scipy.sparse.linalg.expm_multiply (traceA = 0), for CI spaces where the
    dense T of :413-469 is too slow to build (minutes at (8e,8o))."""
    from . import cscan

    hs = cscan.HashedSpace(ci)
    Top = get_ucc_T(thetas, layout)
    d = len(ci.idx2det)

    def mv(v):
        return cscan.propagate_state_c([Top], np.ascontiguousarray(v, dtype=np.float64).ravel(), hs,
                                       nthreads=nthreads)

    A = ssl.LinearOperator((d, d), matvec=mv, rmatvec=lambda v: -mv(v), dtype=np.float64)
    return ssl.expm_multiply(A, np.asarray(state, dtype=np.float64), traceA=0.0)


def gradient_forward_fd(H_mat, csf, ci, thetas, layout):
    """Forward-difference gradient of ucc_wavefunction.py:1069-1097 (step sqrt(eps))."""
    eps = np.finfo(np.float64).eps
    h = eps ** 0.5
    Tmat = operator_matrix(get_ucc_T(thetas, layout), ci)
    psi = ssl.expm_multiply(Tmat, csf, traceA=0.0)
    Hpsi = H_mat @ psi
    E = float(psi @ Hpsi)
    g = np.zeros(len(thetas))
    for i in range(len(thetas)):
        step = np.zeros(len(thetas))
        step[i] = h
        Ti = operator_matrix(get_ucc_T(step, layout), ci)
        psi_p = ssl.expm_multiply(Tmat + Ti, csf, traceA=0.0)
        g[i] = 2 * (psi_p @ Hpsi - E) / h
    return E, g
