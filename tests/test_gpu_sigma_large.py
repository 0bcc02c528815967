"""sigma = H psi at the headline sizes against the oracle's C scan.

The (12e,12o) workload's sigma runs the same kernels as every energy and
gradient of the headline benchmark; here it is compared, on random vectors,
with the oracle's restatement of the reference's determinant scan
(oracle/csrc/scan.c, apply_operator_threaded, operator_state_algebra.py:139-219)
run on all host threads, at (10e,10o), (11e,11o) (6 alpha, 5 beta) and
(12e,12o), and at 13 and 14 orbitals with 4-6 electrons, with synthetic
rank-2n Hamiltonians.  Tolerance 1e-11 relative to
max|sigma| element-wise, energies 1e-10.

The tUPS (12e,12o) 4-layer state is also checked through an independent
route: E from the 1-/2-RDMs (k_gram_mfma, sq_rdm12) contracted with h and g on
the host equals the sigma energy (1e-10), tr Gamma1 = 12 and
sum_r Gamma2[p,q,r,r] = 11 Gamma1[p,q].
"""

from __future__ import annotations

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


# 13 and 14 orbitals with few electrons: the row-sigma kernel shapes of the
# (13e,13o) and (14e,14o) spaces (they depend on the orbital count only) on
# spaces the scan finishes in seconds
@pytest.mark.parametrize("n,na,nb", [(10, 5, 5), (11, 6, 5), (12, 6, 6), (13, 2, 2), (13, 3, 2), (14, 2, 2), (14, 3, 3)])
def test_sigma_against_oracle_scan(n, na, nb):
    from oracle import cscan
    from oracle.ci import get_indexing
    from oracle.fermion import hamiltonian_active_fast
    from slowquant_amd import engine as E
    from slowquant_amd.synthetic import synthetic_problem

    h, g, _ = synthetic_problem(n, 2 * n, 1, 0.05, 100 + n)
    ci = get_indexing(0, n, 0, na, nb)
    sp = E.Space(n, na, nb)
    assert sp.dim == len(ci.idx2det)
    H = E.Hamiltonian(sp, h, g, e_core=0.5)
    rng = np.random.default_rng(n)
    psi = rng.normal(size=sp.dim)
    psi /= np.linalg.norm(psi)
    sig, e = H.sigma(sp.from_reference(psi))
    got = sp.to_reference_numpy(sig)
    print(f"sigma at ({n},{na},{nb}): oracle scan on {_threads()} threads", flush=True)
    hs = cscan.HashedSpace(ci)
    expect = cscan.propagate_state_c([hamiltonian_active_fast(h, g, n)], psi, hs, nthreads=_threads()) + 0.5 * psi
    scale = max(1.0, float(np.abs(expect).max()))
    assert np.abs(got - expect).max() < 1e-11 * scale
    assert abs(e - float(psi @ expect)) < 1e-10


def test_tups12_energy_from_rdms_equals_sigma_energy():
    from slowquant_amd import engine as E
    from slowquant_amd.synthetic import synthetic_problem
    from slowquant_amd.util import UpsStructure

    lay = UpsStructure()
    lay.create_tiled(12, {"n_layers": 4, "do_tups": True})
    h, g, theta = synthetic_problem(12, 24, lay.n_params, 0.05, 0)
    sp = E.Space(12, 6, 6)
    prog = E.Program(sp, lay.excitation_operator_type, lay.excitation_indices)
    H = E.Hamiltonian(sp, h, g)
    psi = prog.state_prep(theta, E.hf_state(sp))
    _, e_sigma = H.sigma(psi)
    e_adj, _ = prog.energy_grad(H, theta, E.hf_state(sp), "adjoint")
    r1, r2 = E.rdm12(sp, psi)
    e_rdm = float(np.einsum("pq,pq->", h, r1) + 0.5 * np.einsum("pqrs,pqrs->", g, r2))
    assert abs(e_rdm - e_sigma) < 1e-10
    assert abs(e_adj - e_sigma) < 1e-10
    assert abs(np.trace(r1) - 12.0) < 1e-10
    assert np.abs(np.einsum("pqrr->pq", r2) - 11.0 * r1).max() < 1e-9
    assert np.abs(r1 - r1.T).max() < 1e-12


@pytest.mark.parametrize("n,na,nb", [(13, 2, 2), (14, 2, 2), (14, 3, 2)])
def test_tups_13_14_orbitals_few_electrons_against_oracle(n, na, nb):
    """tUPS (one layer) at 13 and 14 orbitals with 4-5 electrons: the brick
    stages over 13-14 orbitals, the sigma and the reverse sweep against the
    oracle: the state (1e-12), the energy (1e-10) and the adjoint gradient
    against the oracle's product-rule sweep (ups_wavefunction.py:1090-1141,
    1e-10)."""
    from oracle.apply import expectation_value
    from oracle.ci import get_indexing
    from oracle.fermion import hamiltonian_active_fast
    from oracle.ups import construct_ups_state, gradient_reference
    from slowquant_amd import engine as E
    from slowquant_amd.synthetic import synthetic_hamiltonian
    from slowquant_amd.util import UpsStructure

    lay = UpsStructure()
    lay.create_tiled(n, {"n_layers": 1, "do_tups": True})
    ci = get_indexing(0, n, 0, na, nb)
    theta = np.random.default_rng(n + na).normal(0.0, 0.2, lay.n_params)
    csf = np.zeros(len(ci.idx2det))
    csf[0] = 1.0
    h, g = synthetic_hamiltonian(n, 2 * n, np.random.default_rng(n))
    sp = E.Space(n, na, nb)
    prog = E.Program(sp, lay.excitation_operator_type, lay.excitation_indices)
    H = E.Hamiltonian(sp, h, g)
    psi_ref = construct_ups_state(csf, ci, theta, lay)
    psi = prog.state_prep(theta, sp.from_reference(csf))
    assert np.abs(sp.to_reference_numpy(psi) - psi_ref).max() < 1e-12
    Ho = hamiltonian_active_fast(h, g, n)
    e_ref = expectation_value(psi_ref, [Ho], psi_ref, ci, do_folding=False)
    e, grad = prog.energy_grad(H, theta, sp.from_reference(csf))
    assert abs(e - e_ref) < 1e-10
    assert np.abs(grad - gradient_reference(Ho, csf, ci, theta, lay)).max() < 1e-10
