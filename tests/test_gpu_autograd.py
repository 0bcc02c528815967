"""The adjoint backward behind torch.autograd.Function.

(12e,12o) tUPS with 4 layers (132 parameters, CI dimension 853,776) on the
seeded synthetic Hamiltonian (rank 24, variance 0.3^2/12, theta ~ N(0, 0.05^2),
seed 0): the autograd gradient is checked against central finite
differences of the engine energy (h = 1e-5, tolerance 1e-8), the reference's
product-rule sweep and the plain per-parameter propagation (tolerance 1e-10).
A (6e,6o) case is also checked against the oracle.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(n, layers, rank, sigma, seed):
    from slowquant_amd import engine as E
    from slowquant_amd.synthetic import synthetic_problem
    from slowquant_amd.util import UpsStructure

    lay = UpsStructure()
    lay.create_tiled(n, {"n_layers": layers, "do_tups": True})
    h, g, theta = synthetic_problem(n, rank, lay.n_params, sigma, seed)
    sp = E.Space(n, n // 2, n // 2)
    prog = E.Program(sp, lay.excitation_operator_type, lay.excitation_indices)
    H = E.Hamiltonian(sp, h, g)
    return lay, h, g, theta, sp, prog, H, E.hf_state(sp)


def _energy(prog, H, th, psi0):
    _, e = H.sigma(prog.state_prep(th, psi0))
    return e


def test_autograd_12o_against_fd_and_sweeps():
    import torch

    from slowquant_amd.autograd import ups_energy

    lay, h, g, theta, sp, prog, H, psi0 = _problem(12, 4, 24, 0.05, 0)
    assert sp.dim == 853776 and lay.n_params == 132
    th = torch.tensor(theta, dtype=torch.float64, device="cuda", requires_grad=True)
    e = ups_energy(th, prog, H, psi0)
    e.backward()
    g_auto = th.grad.cpu().numpy()
    e_ref, g_ref = prog.energy_grad(H, theta, psi0, "reference")
    assert abs(float(e) - e_ref) < 1e-10
    assert np.abs(g_auto - g_ref).max() < 1e-10
    _, g_plain = prog.energy_grad(H, theta, psi0, "plain")
    assert np.abs(g_auto - g_plain).max() < 1e-10
    hstep = 1e-5
    g_fd = np.zeros_like(theta)
    for i in range(len(theta)):
        tp = theta.copy()
        tm = theta.copy()
        tp[i] += hstep
        tm[i] -= hstep
        g_fd[i] = (_energy(prog, H, tp, psi0) - _energy(prog, H, tm, psi0)) / (2 * hstep)
    assert np.abs(g_auto - g_fd).max() < 1e-8


def test_autograd_6o_against_oracle_and_chain_rule():
    import torch

    from oracle import fermion as F
    from oracle.apply import expectation_value
    from oracle.ci import get_indexing
    from oracle.ups import construct_ups_state, gradient_reference
    from slowquant_amd.autograd import ups_energy

    lay, h, g, theta, sp, prog, H, psi0 = _problem(6, 2, 12, 0.1, 3)
    th = torch.tensor(theta, dtype=torch.float64, device="cuda", requires_grad=True)
    e = ups_energy(th, prog, H, psi0)
    (3.0 * e).backward()  # chain rule through grad_out
    Hf = F.hamiltonian_0i_0a(h, g, 0, 6)
    ci = get_indexing(0, 6, 0, 3, 3)
    csf = np.zeros(len(ci.idx2det))
    csf[0] = 1.0
    psi = construct_ups_state(csf, ci, theta, lay)
    assert abs(float(e) - expectation_value(psi, [Hf], psi, ci, do_folding=False)) < 1e-10
    g_ref = gradient_reference(Hf, csf, ci, theta, lay)
    assert np.abs(th.grad.cpu().numpy() - 3.0 * g_ref).max() < 1e-10
    with pytest.raises(ValueError):
        ups_energy(th[:-1], prog, H, psi0)


def test_14o_tups6_energy_gradient_rdms():
    """tUPS with 6 layers on the synthetic (14e,14o) Hamiltonian (rank 28,
    variance 0.3^2/14, theta ~ N(0, 0.03^2), seed 0; CI dimension 11,778,624,
    234 parameters): adjoint gradient against the reference sweep (1e-10) and
    central differences on a few parameters (h = 1e-5, 1e-8); RDM identities
    (trace, partial trace, E(Gamma) = <H>, 1e-10)."""
    from slowquant_amd import engine as E

    lay, h, g, theta, sp, prog, H, psi0 = _problem(14, 6, 28, 0.03, 0)
    assert sp.dim == 11778624 and lay.n_params == 234
    e_a, g_a = prog.energy_grad(H, theta, psi0, "adjoint")
    e_r, g_r = prog.energy_grad(H, theta, psi0, "reference")
    assert abs(e_a - e_r) < 1e-10
    assert np.abs(g_a - g_r).max() < 1e-10
    hstep = 1e-5
    for i in (0, 117, 233):
        tp = theta.copy()
        tm = theta.copy()
        tp[i] += hstep
        tm[i] -= hstep
        fd = (_energy(prog, H, tp, psi0) - _energy(prog, H, tm, psi0)) / (2 * hstep)
        assert abs(fd - g_a[i]) < 1e-8
    psi = prog.state_prep(theta, psi0)
    r1, r2 = E.rdm12(sp, psi)
    assert abs(np.trace(r1) - 14.0) < 1e-10
    assert np.abs(np.einsum("pqrr->pq", r2) - 13.0 * r1).max() < 1e-9
    e_rdm = np.einsum("pq,pq->", h, r1) + 0.5 * np.einsum("pqrs,pqrs->", g, r2)
    assert abs(e_rdm - e_a) < 1e-10


def test_14o_large_tile_side_launch_matches_single_launch(monkeypatch):
    """At (14e,14o) the few tiles above the tile target run as a side launch
    beside the main stage launch; with SQ_TILE_SPLIT=0 every tile runs in one
    launch.  Both give the same state, energy and gradient (the tiles and
    their arithmetic are the same; 1e-13)."""
    from slowquant_amd import engine as E

    lay, h, g, theta, sp, prog, H, psi0 = _problem(14, 6, 28, 0.03, 0)
    a = prog.state_prep(theta, psi0)  # builds the stage plan with the side launches
    monkeypatch.setenv("SQ_TILE_SPLIT", "0")
    one = E.Program(sp, lay.excitation_operator_type, lay.excitation_indices)
    b = one.state_prep(theta, psi0)  # plan built with one launch per stage
    assert (a - b).abs().max().item() < 1e-13
    e_s, g_s = prog.energy_grad(H, theta, psi0, "adjoint")
    e_o, g_o = one.energy_grad(H, theta, psi0, "adjoint")
    assert abs(e_s - e_o) < 1e-13
    assert np.abs(g_s - g_o).max() < 1e-13
