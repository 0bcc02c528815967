"""GPU parity of the HIP engine against the CPU oracle (C ABI through ctypes).

Tolerances: states 1e-12 max-abs, energies 1e-10, gradients 1e-8 (FP64
throughout); integer tables bit-exact.
"""

from __future__ import annotations

import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _engine():
    from slowquant_amd import engine

    return engine


def _rand_state(d, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=d)
    return v / np.linalg.norm(v)


def _torch():
    import torch

    return torch


@pytest.mark.parametrize("n,na,nb", [(1, 1, 0), (2, 1, 1), (4, 2, 2), (5, 3, 2), (6, 3, 3), (8, 4, 4), (9, 3, 5)])
def test_space_tables_bit_exact(n, na, nb):
    from oracle.ci import get_indexing, spin_strings

    E = _engine()
    sp = E.Space(n, na, nb)
    a, b = sp.strings()
    ref_a = [sum(1 << p for p in c) for c in spin_strings(n, na)]
    ref_b = [sum(1 << p for p in c) for c in spin_strings(n, nb)]
    assert a.tolist() == ref_a
    assert b.tolist() == ref_b
    ci = get_indexing(0, n, 0, na, nb)
    assert np.array_equal(sp.idx2det(), ci.idx2det)


def test_phase_roundtrip_and_sign():
    torch = _torch()
    E = _engine()
    sp = E.Space(6, 3, 3)
    x = torch.tensor(_rand_state(sp.dim, 1), device="cuda")
    y = sp.to_reference(sp.from_reference(x))
    assert torch.equal(x, y)
    # HF determinant (index 0): phase (-1)^{k(k-1)/2} with k = 3 -> -1
    e0 = torch.zeros(sp.dim, dtype=torch.float64, device="cuda")
    e0[0] = 1.0
    assert float(sp.from_reference(e0)[0]) == -1.0


def test_mfma_gemm_layout():
    torch = _torch()
    import ctypes

    from slowquant_amd import _native as N

    for npair, K4, kpad, dim in [(10, 12, 16, 37), (78, 80, 80, 1000), (105, 108, 112, 517)]:
        g = torch.Generator().manual_seed(npair)
        At = torch.randn(K4, kpad, dtype=torch.float64, generator=g).cuda()
        D = torch.randn(K4, dim, dtype=torch.float64, generator=g).cuda()
        G = torch.full((npair, dim), float("nan"), dtype=torch.float64, device="cuda")
        N.check(N.lib().sq_debug_gemm(At.data_ptr(), D.data_ptr(), G.data_ptr(), npair, K4, kpad, dim,
                                      torch.cuda.current_stream().cuda_stream))
        ref = (At.cpu().T @ D.cpu())[:npair]
        assert torch.allclose(G.cpu(), ref, rtol=1e-12, atol=1e-12)
    del ctypes


GATES_44 = [
    ("single", (0, 4)),
    ("single", (1, 7)),
    ("single", (2, 6)),
    ("double", (0, 1, 2, 3)),
    ("double", (0, 3, 4, 7)),
    ("double", (2, 1, 6, 5)),
    ("double", (0, 2, 4, 6)),
    ("triple", (0, 1, 2, 4, 5, 6)),
    ("quadruple", (0, 1, 2, 3, 4, 5, 6, 7)),
    ("sa_single", (0, 2)),
    ("sa_single", (1, 2)),
    ("sa_single", (0, 3)),
    ("sa_double_1", (0, 0, 2, 2)),
    ("sa_double_2", (0, 0, 2, 3)),
    ("sa_double_3", (0, 1, 2, 2)),
    ("sa_double_4", (0, 1, 2, 3)),
    ("sa_double_5", (0, 1, 2, 3)),
    ("sa_double_4", (1, 0, 3, 2)),
]


@pytest.mark.parametrize("gate", GATES_44)
@pytest.mark.parametrize("theta", [0.37, -1.3])
def test_single_gate_parity(gate, theta):
    torch = _torch()
    from oracle.ci import get_indexing
    from oracle.ups import apply_gate, get_grad_action

    E = _engine()
    t, idx = gate
    ci = get_indexing(0, 4, 0, 2, 2)
    sp = E.Space(4, 2, 2)
    prog = E.Program(sp, [t], [idx])
    psi_ref = _rand_state(sp.dim, 7)
    expect = apply_gate(psi_ref, t, idx, theta, ci)
    psi = sp.from_reference(psi_ref)
    prog.apply_gate(0, [theta], psi)
    got = sp.to_reference_numpy(psi)
    assert np.abs(got - expect).max() < 1e-12

    class L:
        excitation_operator_type = [t]
        excitation_indices = [idx]

    expect_t = get_grad_action(psi_ref, 0, ci, L)
    got_t = sp.to_reference_numpy(prog.grad_action(0, sp.from_reference(psi_ref)))
    assert np.abs(got_t - expect_t).max() < 1e-12
    del torch


def _tups(n, layers):
    from slowquant_amd.util import UpsStructure

    s = UpsStructure()
    s.create_tiled(n, {"n_layers": layers, "do_tups": True})
    return s


def _hamiltonian(n, rank, seed):
    from slowquant_amd.synthetic import synthetic_hamiltonian

    return synthetic_hamiltonian(n, rank, np.random.default_rng(seed))


@pytest.mark.parametrize("n,layers", [(4, 1), (6, 2)])
def test_state_prep_parity(n, layers):
    from oracle.ci import get_indexing
    from oracle.ups import construct_ups_state

    E = _engine()
    lay = _tups(n, layers)
    rng = np.random.default_rng(3)
    theta = rng.normal(0, 0.3, lay.n_params)
    ci = get_indexing(0, n, 0, n // 2, n // 2)
    sp = E.Space(n, n // 2, n // 2)
    prog = E.Program(sp, lay.excitation_operator_type, lay.excitation_indices)
    csf = np.zeros(sp.dim)
    csf[0] = 1.0
    expect = construct_ups_state(csf, ci, theta, lay)
    got = sp.to_reference_numpy(prog.state_prep(theta, sp.from_reference(csf)))
    assert np.abs(got - expect).max() < 1e-12
    # dagger undoes the state preparation
    back = sp.to_reference_numpy(prog.state_prep(theta, sp.from_reference(expect), dagger=True))
    assert np.abs(back - csf).max() < 1e-12
    expect_d = construct_ups_state(csf, ci, theta, lay, dagger=True)
    got_d = sp.to_reference_numpy(prog.state_prep(theta, sp.from_reference(csf), dagger=True))
    assert np.abs(got_d - expect_d).max() < 1e-12


def _oracle_H(h, g, n):
    from oracle.fermion import hamiltonian_active_fast

    return hamiltonian_active_fast(h, g, n)


@pytest.mark.parametrize("n,na,nb", [(4, 2, 2), (6, 3, 3), (6, 2, 4), (8, 4, 4)])
def test_sigma_parity(n, na, nb):
    from oracle.apply import propagate_state
    from oracle.ci import get_indexing

    E = _engine()
    h, g = _hamiltonian(n, 2 * n, 5)
    ci = get_indexing(0, n, 0, na, nb)
    sp = E.Space(n, na, nb)
    H = E.Hamiltonian(sp, h, g, e_core=0.25)
    psi_ref = _rand_state(sp.dim, 11)
    expect = propagate_state([_oracle_H(h, g, n)], psi_ref, ci, do_folding=False) + 0.25 * psi_ref
    sig, e = H.sigma(sp.from_reference(psi_ref))
    got = sp.to_reference_numpy(sig)
    assert np.abs(got - expect).max() < 1e-11
    assert abs(e - float(psi_ref @ expect)) < 1e-10


@pytest.mark.parametrize("n,na,nb", [(2, 2, 2), (3, 1, 0), (4, 2, 0)])
def test_degenerate_spaces(n, na, nb):
    """Spaces with one determinant (both spins full) or one beta string (no
    beta electrons): tUPS state, energy, adjoint and reference-sweep
    gradients and RDMs against the oracle; with one determinant every gate
    acts trivially (gradient 0)."""
    from oracle.apply import expectation_value
    from oracle.ci import get_indexing
    from oracle.rdm import rdm1, rdm2
    from oracle.ups import construct_ups_state, gradient_reference

    E = _engine()
    lay = _tups(n, 1)
    h, g = _hamiltonian(n, 2 * n, 21)
    theta = np.random.default_rng(6).normal(0, 0.3, lay.n_params)
    ci = get_indexing(0, n, 0, na, nb)
    Hop = _oracle_H(h, g, n)
    csf = np.zeros(len(ci.idx2det))
    csf[0] = 1.0
    psi_ref = construct_ups_state(csf, ci, theta, lay)
    e_ref = expectation_value(psi_ref, [Hop], psi_ref, ci, do_folding=False)
    g_ref = gradient_reference(Hop, csf, ci, theta, lay)
    sp = E.Space(n, na, nb)
    assert sp.dim == len(ci.idx2det)
    prog = E.Program(sp, lay.excitation_operator_type, lay.excitation_indices)
    H = E.Hamiltonian(sp, h, g)
    psi0 = sp.from_reference(csf)
    assert np.abs(sp.to_reference_numpy(prog.state_prep(theta, psi0)) - psi_ref).max() < 1e-12
    for method in ("adjoint", "reference"):
        e, grad = prog.energy_grad(H, theta, psi0, method)
        assert abs(e - e_ref) < 1e-10, method
        assert np.abs(grad - g_ref).max() < 1e-10, method
    if sp.dim == 1:
        assert np.abs(g_ref).max() < 1e-14
    r1, r2 = E.rdm12(sp, sp.from_reference(psi_ref))
    e1 = rdm1(psi_ref, ci)
    assert np.abs(r1 - e1).max() < 1e-12
    assert np.abs(r2 - rdm2(psi_ref, ci, e1)).max() < 1e-12


def test_empty_program():
    """A program with no gates: the state is psi0 and the gradient is empty."""
    E = _engine()
    h, g = _hamiltonian(4, 8, 3)
    sp = E.Space(4, 2, 2)
    H = E.Hamiltonian(sp, h, g)
    prog = E.Program(sp, [], [])
    psi0 = sp.from_reference(_rand_state(sp.dim, 2))
    out = prog.state_prep(np.zeros(0), psi0)
    assert np.array_equal(out.cpu().numpy(), psi0.cpu().numpy())
    e, grad = prog.energy_grad(H, np.zeros(0), psi0, "adjoint")
    _sig, e_ref = H.sigma(psi0)
    assert abs(e - e_ref) < 1e-12 and grad.shape == (0,)


@pytest.mark.parametrize("n,layers", [(4, 1), (6, 1)])
def test_gradients_parity(n, layers):
    from oracle.ci import get_indexing
    from oracle.ups import construct_ups_state, gradient_reference

    E = _engine()
    lay = _tups(n, layers)
    h, g = _hamiltonian(n, 2 * n, 9)
    theta = np.random.default_rng(4).normal(0, 0.2, lay.n_params)
    ci = get_indexing(0, n, 0, n // 2, n // 2)
    Hop = _oracle_H(h, g, n)
    csf = np.zeros(len(ci.idx2det))
    csf[0] = 1.0
    g_ref = gradient_reference(Hop, csf, ci, theta, lay)
    from oracle.apply import expectation_value

    psi = construct_ups_state(csf, ci, theta, lay)
    e_ref = expectation_value(psi, [Hop], psi, ci, do_folding=False)
    sp = E.Space(n, n // 2, n // 2)
    prog = E.Program(sp, lay.excitation_operator_type, lay.excitation_indices)
    H = E.Hamiltonian(sp, h, g)
    psi0 = sp.from_reference(csf)
    for method in ("adjoint", "plain", "reference"):
        e, grad = prog.energy_grad(H, theta, psi0, method)
        assert abs(e - e_ref) < 1e-10, method
        assert np.abs(grad - g_ref).max() < 1e-10, method


def test_fucc_sa_doubles_parity():
    from oracle.ci import get_indexing
    from oracle.ups import construct_ups_state, gradient_reference

    from slowquant_amd.util import UpsStructure

    E = _engine()
    lay = UpsStructure()
    lay.create_fUCC([0, 1], [2, 3], [0, 1, 2, 3], [4, 5, 6, 7], 4, {"n_layers": 1, "SAS": True, "SAD": True})
    theta = np.random.default_rng(8).normal(0, 0.3, lay.n_params)
    ci = get_indexing(0, 4, 0, 2, 2)
    csf = np.zeros(36)
    csf[0] = 1.0
    expect = construct_ups_state(csf, ci, theta, lay)
    sp = E.Space(4, 2, 2)
    prog = E.Program(sp, lay.excitation_operator_type, lay.excitation_indices)
    got = sp.to_reference_numpy(prog.state_prep(theta, sp.from_reference(csf)))
    assert np.abs(got - expect).max() < 1e-12
    h, g = _hamiltonian(4, 8, 2)
    H = E.Hamiltonian(sp, h, g)
    g_ref = gradient_reference(_oracle_H(h, g, 4), csf, ci, theta, lay)
    for method in ("adjoint", "plain", "reference"):
        _e, grad = prog.energy_grad(H, theta, sp.from_reference(csf), method)
        assert np.abs(grad - g_ref).max() < 1e-10, method


@pytest.mark.parametrize(
    "n,na,nb",
    [(4, 2, 2), (6, 3, 3), (5, 2, 3), (8, 4, 4), (9, 4, 4), (10, 5, 4), (11, 2, 3), (12, 2, 2), (13, 1, 2), (14, 2, 1)],
)
def test_rdm_parity(n, na, nb):
    """RDMs against the oracle's element-by-element expectation values.  The
    Gram kernel's tile layout depends on the orbital count only (pairs
    p < q, their transposes, the diagonal), so few electrons keep the oracle
    fast at 11-14 orbitals, where the layout has 4-6 tiles per block and the
    transposition symmetry rebuilds most of the matrix."""
    from oracle.ci import get_indexing
    from oracle.rdm import rdm1, rdm2

    E = _engine()
    ci = get_indexing(0, n, 0, na, nb)
    sp = E.Space(n, na, nb)
    psi_ref = _rand_state(sp.dim, 13)
    r1, r2 = E.rdm12(sp, sp.from_reference(psi_ref))
    e1 = rdm1(psi_ref, ci)
    e2 = rdm2(psi_ref, ci, e1)
    assert np.abs(r1 - e1).max() < 1e-12
    assert np.abs(r2 - e2).max() < 1e-12


@pytest.mark.parametrize(
    "n,na,nb,opts",
    [
        (6, 3, 3, {"n_layers": 2, "do_tups": True}),
        (5, 2, 2, {"n_layers": 2, "do_tups": True}),
        (8, 4, 4, {"n_layers": 2, "do_qnp": True}),
        (8, 3, 5, {"n_layers": 2, "do_tups": True}),
        (9, 5, 3, {"n_layers": 1, "do_qnp": True}),
        (10, 5, 5, {"n_layers": 1, "do_tups": True, "skip_last_singles": True}),
        (14, 7, 7, {"n_layers": 1, "do_tups": True}),
    ],
)
def test_tiled_stages_match_per_gate_kernels(n, na, nb, opts, monkeypatch):
    """Fused brick-column passes (default) against the per-gate rotation kernels
    (SQ_TILES=0) for state preparation, dagger and all three gradients,
    including spaces with unequal alpha and beta electron counts."""
    from slowquant_amd.synthetic import synthetic_problem
    from slowquant_amd.util import UpsStructure

    E = _engine()
    lay = UpsStructure()
    lay.create_tiled(n, dict(opts))
    h, g, theta = synthetic_problem(n, n, lay.n_params, 0.2, 5)
    sp = E.Space(n, na, nb)
    tiled = E.Program(sp, lay.excitation_operator_type, lay.excitation_indices)
    monkeypatch.setenv("SQ_TILES", "0")
    plain_prog = E.Program(sp, lay.excitation_operator_type, lay.excitation_indices)
    psi0 = E.hf_state(sp)
    a = plain_prog.state_prep(theta, psi0).cpu().numpy()
    monkeypatch.setenv("SQ_TILES", "1")
    b = tiled.state_prep(theta, psi0).cpu().numpy()
    assert np.abs(a - b).max() < 1e-12
    back = tiled.state_prep(theta, tiled.state_prep(theta, psi0), dagger=True).cpu().numpy()
    assert np.abs(back - psi0.cpu().numpy()).max() < 1e-12
    if n <= 10:
        H = E.Hamiltonian(sp, h, g)
        e0, g0 = plain_prog.energy_grad(H, theta, psi0, "adjoint")
        for method in ("adjoint", "plain", "reference"):
            e1, g1 = tiled.energy_grad(H, theta, psi0, method)
            assert abs(e1 - e0) < 1e-10, method
            assert np.abs(g1 - g0).max() < 1e-10, method


@pytest.mark.parametrize("n,layers", [(8, 2), (10, 1)])
def test_tile_dispatch_order_does_not_change_results(n, layers, monkeypatch):
    """The XCD-aware tile dispatch order (default) and the plain largest-first
    order (SQ_TILE_XCD=0) give bit-identical states, energies and gradients:
    every tile writes its own amplitudes and its own gradient partials, which
    are summed in tile order whatever order the tiles ran in."""
    from slowquant_amd.synthetic import synthetic_problem

    from slowquant_amd import _native as N

    E = _engine()
    lay = _tups(n, layers)
    h, g, theta = synthetic_problem(n, n, lay.n_params, 0.2, 9)
    sp = E.Space(n, n // 2, n // 2)
    H = E.Hamiltonian(sp, h, g)
    psi0 = E.hf_state(sp)
    out = []
    N.check(N.lib().sq_sigma_configure(sp.handle, 0, 0, 1))  # sigma with fixed-order sums
    try:
        for flag in ("0", "1"):
            monkeypatch.setenv("SQ_TILE_XCD", flag)
            prog = E.Program(sp, lay.excitation_operator_type, lay.excitation_indices)
            psi = prog.state_prep(theta, psi0).cpu().numpy()  # the tile plan is built here
            e, grad = prog.energy_grad(H, theta, psi0, "adjoint")
            out.append((psi, e, grad))
    finally:
        N.check(N.lib().sq_sigma_configure(sp.handle, 0, 0, -1))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2])
    # likewise the RDM chunk order (each chunk writes its own partial)
    rd = []
    psi_t = prog.state_prep(theta, psi0)
    for flag in ("0", "1"):
        monkeypatch.setenv("SQ_GRAM_XCD", flag)
        rd.append(E.rdm12(sp, psi_t))
    assert np.array_equal(rd[0][0], rd[1][0]) and np.array_equal(rd[0][1], rd[1][1])


def test_back_to_back_theta_uploads():
    """Host theta goes through one pinned staging buffer per program: calls
    issued without a host synchronization in between must each see their own
    theta."""
    torch = _torch()
    E = _engine()
    lay = _tups(6, 2)
    sp = E.Space(6, 3, 3)
    prog = E.Program(sp, lay.excitation_operator_type, lay.excitation_indices)
    psi0 = E.hf_state(sp)
    rng = np.random.default_rng(21)
    thetas = [rng.normal(0, 0.3, lay.n_params) for _ in range(4)]
    outs = [prog.state_prep(th, psi0) for th in thetas]  # no synchronization between calls
    torch.cuda.synchronize()
    for th, out in zip(thetas, outs):
        ref = prog.state_prep(torch.tensor(th, dtype=torch.float64, device="cuda"), psi0)
        assert torch.equal(out, ref)
