"""CPU oracle for the SlowQuant state-vector path -- TEST INFRASTRUCTURE ONLY.

This package is a from-scratch CPU restatement of the reference algorithms that
the HIP engine in ``slowquant_amd`` replaces.  It exists to check the engine and
to time the reference algorithm on host cores.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it, and only as the checker; the product path never routes through it.

Modules
-------
ci          determinant addressing (ci_spaces.py:56-116)
fermion     normal-ordered fermionic operator algebra and folding
            (fermionic_operator.py:27-321,379-471; operators.py:8-529)
apply       the determinant scan of one operator string and propagate_state /
            expectation_value (operator_state_algebra.py:33-219,472-630,784-824)
ups         gate closed forms, construct_ups_state, propagate_unitary,
            get_grad_action and the three gradient algorithms
            (operator_state_algebra.py:963-1412,1867-2309,2757-2865;
            ups_wavefunction.py:1060-1142)
ucc         get_ucc_T / build_operator_matrix / expm_multiply
            (operator_state_algebra.py:413-469,870-960)
rdm         rdm1/rdm2 by expectation values (ups_wavefunction.py:409-476)
dense       brute-force dense second quantisation, independent of ``fermion``
cscan       ctypes wrapper of the C restatement of the Numba scan kernel

Pinning status (what fixes each function independently of its own formula)
--------------------------------------------------------------------------
ci.get_indexing        brute force over all bit strings (tests/test_oracle.py)
fermion normal order   dense matrices of raw vs normal-ordered products;
                       golden G2_sa strings from the reference (tests/golden)
apply scan             dense brute-force matrices (oracle.dense); C scan agrees
ups gates              scipy.linalg.expm of every generator, all 13 gate types
ups gradients          central finite differences; the three algorithms agree
hamiltonian + folding  H4 rectangle STO-3G FCI known answer -3.714153922167;
                       full-space brute force vs folded active space
rdm                    trace identities, E(rdm) == <H>, partial traces
ucc                    scipy.linalg.expm vs expm_multiply
No function of this package is left "parity unpinned".
"""
