"""GPU checks of the generic operator path (propagate_state /
expectation_value with FermionicOperators, operator_state_algebra.py:472-824,
as linear response calls it) beyond (4e,4o): the first application of a
compiled operator scans its terms per determinant, later ones run through the
operator's sparsity pattern; both against the oracle, and the pattern path
against the independent sigma kernel."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _space(n, na, nb):
    from oracle.ci import get_indexing as o_idx

    from slowquant_amd.ci_spaces import get_indexing

    return get_indexing(0, n, 0, na, nb), o_idx(0, n, 0, na, nb)


def _state(dim, seed):
    v = np.random.default_rng(seed).normal(size=dim)
    return v / np.linalg.norm(v)


def test_hamiltonian_operator_6o_scan_and_pattern_against_oracle():
    """H = hamiltonian_0i_0a of a synthetic (6e,6o) h/g as a FermionicOperator,
    applied three times (scan, then pattern twice): every result equals the
    oracle's string scan to 1e-12, the pattern results equal the scan to
    1e-13, and <psi|H|psi> matches to 1e-12."""
    from oracle import fermion as F
    from oracle.apply import propagate_state as o_prop

    from slowquant_amd import operator_state_algebra as OSA
    from slowquant_amd import operators as OP
    from slowquant_amd.synthetic import synthetic_problem

    n = 6
    h, g, _ = synthetic_problem(n, 12, 1, 0.1, 5)
    ci, oci = _space(n, 3, 3)
    psi = _state(400, 1)
    H = OP.hamiltonian_0i_0a(h, g, 0, n)
    ref = o_prop([F.hamiltonian_0i_0a(h, g, 0, n)], psi, oci, do_folding=False)
    outs = [OSA.propagate_state([H], psi, ci, do_folding=False) for _ in range(3)]
    for out in outs:
        assert np.abs(out - ref).max() < 1e-12
    assert np.abs(outs[1] - outs[0]).max() < 1e-13
    assert np.abs(outs[2] - outs[1]).max() == 0.0
    e = OSA.expectation_value(psi, [H], psi, ci, do_folding=False)
    assert abs(e - float(psi @ ref)) < 1e-12


@pytest.mark.parametrize("n,na,nb", [(6, 3, 3), (6, 2, 3), (8, 4, 4)])
def test_excitation_operators_repeated_against_oracle(n, na, nb):
    """Linear-response style operators (singlet singles and doubles, plain
    doubles, their adjoints, and "U"/"Ud" sandwiches with a tUPS layout) on
    random states, each applied twice through the module API (term scan,
    then pattern), against the oracle to 1e-12."""
    from oracle import fermion as F
    from oracle.apply import propagate_state as o_prop
    from oracle.ups import construct_ups_state as o_ups

    from slowquant_amd import operator_state_algebra as OSA
    from slowquant_amd import operators as OP
    from slowquant_amd.util import UpsStructure

    from math import comb

    ci, oci = _space(n, na, nb)
    psi = _state(comb(n, na) * comb(n, nb), 2)
    lay = UpsStructure()
    lay.create_tiled(n, {"n_layers": 1, "do_tups": True})
    th = np.random.default_rng(4).normal(0, 0.2, lay.n_params)
    pairs = [
        (OP.G1_sa(0, n - 1), F.G1_sa(0, n - 1)),
        (OP.G2_sa(0, 1, n - 2, n - 1, 1), F.G2_sa(0, 1, n - 2, n - 1, 1)),
        (OP.G2(0, 3, 2 * n - 2, 2 * n - 1), F.G2(0, 3, 2 * n - 2, 2 * n - 1)),
        (OP.Epq(1, n - 2) * OP.Epq(n - 1, 0), F.Epq(1, n - 2) * F.Epq(n - 1, 0)),
    ]
    for op, oop in pairs:
        ref = o_prop([oop], psi, oci)
        ref_d = o_prop([oop.dagger], psi, oci)
        for _ in range(2):
            assert np.abs(OSA.propagate_state([op], psi, ci) - ref).max() < 1e-12
            assert np.abs(OSA.propagate_state([op.dagger], psi, ci) - ref_d).max() < 1e-12
        v = o_ups(psi, oci, th, lay)
        v = o_prop([oop], v, oci)
        v = o_ups(v, oci, th, lay, dagger=True)
        for _ in range(2):
            got = OSA.propagate_state(["Ud", op, "U"], psi, ci, th, lay)
            assert np.abs(got - v).max() < 1e-12


def test_pattern_matches_term_scan_and_sigma_8o(monkeypatch):
    """(8e,8o), CI dimension 4900: H as a FermionicOperator through the term
    scan (SQ_UCC_CSR=0), through its pattern, and the engine's sigma kernel
    (a separate algorithm: alpha-string Hamiltonian GEMM plus the packed
    g contraction) agree to 1e-11; a number-changing operator raises
    KeyError as in the reference."""
    import torch

    from slowquant_amd import engine as E
    from slowquant_amd import operator_state_algebra as OSA
    from slowquant_amd import operators as OP
    from slowquant_amd.synthetic import synthetic_problem

    n = 8
    h, g, _ = synthetic_problem(n, 16, 1, 0.1, 7)
    ci, _ = _space(n, 4, 4)
    psi = _state(4900, 3)
    H = OP.hamiltonian_0i_0a(h, g, 0, n)
    OSA._OPERATORS.clear()
    monkeypatch.setenv("SQ_UCC_CSR", "0")
    scan = [OSA.propagate_state([H], psi, ci, do_folding=False) for _ in range(2)]
    OSA._OPERATORS.clear()
    monkeypatch.delenv("SQ_UCC_CSR")
    pat = [OSA.propagate_state([H], psi, ci, do_folding=False) for _ in range(2)]
    assert np.abs(scan[1] - scan[0]).max() == 0.0
    assert np.abs(pat[1] - scan[0]).max() < 1e-12
    sp = E.space_for(n, 4, 4)
    Hd = E.Hamiltonian(sp, h, g)
    sig, _ = Hd.sigma(sp.from_reference(psi))
    torch.cuda.synchronize()
    sig_ref = sp.to_reference_numpy(sig)
    assert np.abs(sig_ref - pat[1]).max() < 1e-11
    with pytest.raises(KeyError):
        OSA.propagate_state([OP.a_op_spin(0, True)], psi, ci)
