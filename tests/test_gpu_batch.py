"""Batched persistent launch (engine.Batch) against per-instance engine calls
and, for the small spaces, the oracle.

Instances: tUPS with 1-3 layers over (4e,4o), (6e,6o), (8e,8o), (10e,10o)
with seeded synthetic Hamiltonians (rank 2n, variance 0.3^2/n) and
theta ~ N(0, 0.1^2).  Tolerances: energy 1e-10, gradient 1e-10 (analytic vs
analytic); the oracle comparison uses the same bounds.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _specs():
    from slowquant_amd.batch import InstanceSpec

    return [
        InstanceSpec(4, 2, 2, 1, 11),
        InstanceSpec(6, 3, 3, 2, 12),
        InstanceSpec(8, 4, 4, 3, 13),
        InstanceSpec(10, 5, 5, 1, 14),
        InstanceSpec(4, 2, 2, 3, 15),
        InstanceSpec(6, 3, 3, 1, 16),
        InstanceSpec(8, 4, 4, 2, 17),
        InstanceSpec(4, 2, 2, 2, 18),
    ]


def _close(e, g, e_ref, g_ref, tol=1e-10):
    assert np.abs(np.asarray(e) - np.asarray(e_ref)).max() < tol
    for a, b in zip(g, g_ref):
        assert a.shape == b.shape
        assert np.abs(a - b).max() < tol


def test_batch_matches_instance_loop_and_oracle():
    from slowquant_amd import engine as E
    from slowquant_amd.batch import build_instances, instance_problem, loop_energy_grad

    specs = _specs()
    inst, thetas = build_instances(specs)
    e_loop, g_loop = loop_energy_grad(inst, thetas)
    for nwg in (3, 8):  # several instances per workgroup, and one each
        b = E.Batch(inst, nwg)
        e, g = b.run(thetas)
        _close(e, g, e_loop, g_loop)
    # oracle for the (4e,4o) and (6e,6o) instances
    from oracle import fermion as F
    from oracle.apply import expectation_value
    from oracle.ci import get_indexing
    from oracle.ups import construct_ups_state, gradient_reference

    for k, s in enumerate(specs):
        if s.n_orb > 6:
            continue
        layout, h, gg, theta = instance_problem(s)
        Hf = F.hamiltonian_0i_0a(h, gg, 0, s.n_orb)
        ci = get_indexing(0, s.n_orb, 0, s.n_alpha, s.n_beta)
        csf = np.zeros(len(ci.idx2det))
        csf[0] = 1.0
        psi = construct_ups_state(csf, ci, theta, layout)
        e_ref = expectation_value(psi, [Hf], psi, ci, do_folding=False)
        g_ref = gradient_reference(Hf, csf, ci, theta, layout)
        assert abs(e[k] - e_ref) < 1e-10
        assert np.abs(g[k] - g_ref).max() < 1e-10


def test_batch_kernel_shapes():
    """Every autotuner shape (rotation pairs in flight, GEMM column tiles per
    wave) gives the per-instance results, eagerly and through the graph."""
    from slowquant_amd import engine as E
    from slowquant_amd.batch import build_instances, loop_energy_grad

    specs = _specs()
    inst, thetas = build_instances(specs)
    e_loop, g_loop = loop_energy_grad(inst, thetas)
    b = E.Batch(inst, 3)
    for shape in E.SHAPES:
        b.configure(*shape)
        e, g = b.run(thetas)
        _close(e, g, e_loop, g_loop)
        e, g = b.run(thetas, graph=True)
        _close(e, g, e_loop, g_loop)
    with pytest.raises(ValueError):
        b.configure(3, 1)
    with pytest.raises(ValueError):
        b.configure(1, 4)


def test_batch_graph_replays():
    """Replays of the captured graph with new theta and eager work between
    them; every replay equals the eager per-instance result."""
    import torch

    from slowquant_amd import engine as E
    from slowquant_amd.batch import build_instances, loop_energy_grad

    specs = _specs()[:5]
    inst, thetas = build_instances(specs)
    b = E.Batch(inst, 4)
    rng = np.random.default_rng(5)
    for rep in range(3):
        th = [t + rng.normal(0.0, 0.02, t.shape) for t in thetas] if rep else thetas
        e, g = b.run(th, graph=True)
        sp, prog, H = inst[2]
        H.sigma(E.hf_state(sp))  # unrelated eager work on the same stream
        torch.cuda.synchronize()
        e_loop, g_loop = loop_energy_grad(inst, th)
        _close(e, g, e_loop, g_loop)


def test_batch_autotune_and_errors():
    from slowquant_amd import engine as E
    from slowquant_amd.batch import build_instances, loop_energy_grad

    specs = _specs()[:4]
    inst, thetas = build_instances(specs)
    b = E.autotune_batch(inst, np.concatenate(thetas), candidates=(1, 2, 4))
    assert b.tuned_ms > 0
    e, g = b.run(thetas, graph=True)
    e_loop, g_loop = loop_energy_grad(inst, thetas)
    _close(e, g, e_loop, g_loop)
    with pytest.raises(ValueError):
        b.set_theta(np.zeros(b.n_params + 1))
    # polynomial (sa_double_2) gates are not rotations: rejected with ValueError
    sp = E.space_for(4, 2, 2)
    prog = E.Program(sp, ["sa_double_2"], [(0, 1, 2, 3)])
    H = inst[0][2]
    with pytest.raises(ValueError):
        E.Batch([(sp, prog, H)], 1)


def test_batch_start_states():
    """Instances started from given states (random unit vectors, engine
    convention) equal Program.energy_grad from the same states, eagerly and
    over graph replays that see a changed start table; None restores the HF
    start.  Argument errors are raised on the host."""
    import torch

    from slowquant_amd import engine as E
    from slowquant_amd.batch import build_instances, loop_energy_grad

    specs = _specs()[:6]
    inst, thetas = build_instances(specs)
    b = E.Batch(inst, 4)
    gen = torch.Generator().manual_seed(21)

    def unit_states():
        out = []
        for sp, _, _ in inst:
            x = torch.randn(sp.dim, generator=gen, dtype=torch.float64)
            out.append((x / x.norm()).cuda())
        return out

    def expected(starts):
        e, g = [], []
        for (sp, prog, H), th, x in zip(inst, thetas, starts):
            ei, gi = prog.energy_grad(H, th, E.hf_state(sp) if x is None else x)
            e.append(ei)
            g.append(gi)
        return e, g

    starts = unit_states()
    b.set_start_states(starts)
    e, g = b.run(thetas)
    _close(e, g, *expected(starts))
    for rep in range(2):
        e, g = b.run(thetas, graph=True)
        _close(e, g, *expected(starts))
        starts = unit_states()
        b.set_start_states(starts)
        torch.cuda.synchronize()
    mixed = [None if i % 2 else s for i, s in enumerate(starts)]
    b.set_start_states(mixed)
    e, g = b.run(thetas, graph=True)
    _close(e, g, *expected(mixed))
    b.set_start_states([None] * len(inst))
    e, g = b.run(thetas)
    _close(e, g, *loop_energy_grad(inst, thetas))
    with pytest.raises(ValueError):
        b.set_start_states(starts[:-1])
    with pytest.raises(ValueError):
        b.set_start_states([starts[1]] + starts[1:])  # wrong dimension for instance 0
    with pytest.raises(TypeError):
        b.set_start_states([starts[0].cpu()] + starts[1:])


def test_batch512_workload():
    """The 512-instance workload exactly as benchmarked (draw_specs: rng seed 7,
    spaces uniform over (4e,4o)..(10e,10o), 1-3 layers, instance seeds 1000+k,
    rank 2n): autotuned, run through the captured hipGraph, every energy and
    gradient equal to the per-instance loop (1e-10), and the (4e,4o) and (6e,6o)
    instances against the oracle (1e-10)."""
    from oracle import fermion as F
    from oracle.apply import expectation_value
    from oracle.ci import get_indexing
    from oracle.ups import construct_ups_state, gradient_reference
    from slowquant_amd import engine as E
    from slowquant_amd.batch import build_instances, draw_specs, instance_problem, loop_energy_grad

    specs = draw_specs(512)
    assert len(specs) == 512 and {s.n_orb for s in specs} == {4, 6, 8, 10}
    inst, thetas = build_instances(specs)
    b = E.autotune_batch(inst, np.concatenate(thetas))
    e, g = b.run(thetas, graph=True)
    # a second replay gives the same numbers to rounding: the grouped (10e,10o)
    # instances run the row sigma, whose LDS-atomic sums may add in another
    # order (a replay that reused stale or cleared memory would be off by far more)
    e2, g2 = b.run(thetas, graph=True)
    _close(e2, g2, e, g, 1e-12)
    e_loop, g_loop = loop_energy_grad(inst, thetas)
    _close(e, g, e_loop, g_loop)
    print(f"batch512: workgroups {b.n_workgroups}, shape {b.shape}", flush=True)
    for k, s in enumerate(specs):
        if s.n_orb > 6:
            continue
        layout, h, gg, theta = instance_problem(s)
        Hf = F.hamiltonian_0i_0a(h, gg, 0, s.n_orb)
        ci = get_indexing(0, s.n_orb, 0, s.n_alpha, s.n_beta)
        csf = np.zeros(len(ci.idx2det))
        csf[0] = 1.0
        psi = construct_ups_state(csf, ci, theta, layout)
        assert abs(e[k] - expectation_value(psi, [Hf], psi, ci, do_folding=False)) < 1e-10
        assert np.abs(g[k] - gradient_reference(Hf, csf, ci, theta, layout)).max() < 1e-10


def test_batch_groups_match_persistent_launch(monkeypatch):
    """(10e,10o) instances run in groups (tile stages, row sigma and gradient
    kernels with one grid row per instance; three gate layouts, so three
    groups) beside a (6e,6o) instance in the persistent launch: energies and
    gradients equal the all-persistent batch (SQ_BATCH_GROUPS=0) and the
    per-instance loop to 1e-10, across graph replays; after a grouped
    program's tile shape or a grouped space's sigma shape changes, the batch's
    launches raise ValueError instead of using buffers sized for the old
    plan."""
    from slowquant_amd import engine as E
    from slowquant_amd.batch import InstanceSpec, build_instances, loop_energy_grad

    specs = [InstanceSpec(10, 5, 5, L, seed) for L, seed in ((1, 21), (2, 22), (3, 23), (2, 24))]
    specs.append(InstanceSpec(6, 3, 3, 1, 25))
    inst, thetas = build_instances(specs)
    b = E.Batch(inst)
    e, g = b.run(thetas, graph=True)
    e2, g2 = b.run(thetas, graph=True)
    _close(e2, g2, e, g, 1e-12)
    e3, g3 = b.run(thetas, graph=False)
    _close(e3, g3, e, g, 1e-12)
    monkeypatch.setenv("SQ_BATCH_GROUPS", "0")
    b0 = E.Batch(inst)
    e0, g0 = b0.run(thetas, graph=True)
    _close(e, g, e0, g0)
    e_l, g_l = loop_energy_grad(inst, thetas)
    _close(e, g, e_l, g_l)
    inst[0][1].set_tile_shape(1024, 256)
    with pytest.raises(ValueError):
        b.run(thetas, graph=True)
    with pytest.raises(ValueError):
        b.run(thetas, graph=False)
    e4, g4 = b0.run(thetas, graph=True)  # the all-persistent batch does not use tile plans
    _close(e4, g4, e0, g0, 1e-12)
    # the same for the grouped space's row-sigma shape (block count of the
    # grouped sigma launch): refused after a change, accepted once restored
    monkeypatch.delenv("SQ_BATCH_GROUPS")
    sp10 = inst[0][0]
    sp10.configure_sigma(0, 0, -1)  # spaces are shared per process: start from the automatic shape
    b2 = E.Batch(inst)
    _close(*b2.run(thetas, graph=True), e0, g0)
    try:
        sp10.configure_sigma(4, 1, -1)
        with pytest.raises(ValueError):
            b2.run(thetas, graph=True)
        with pytest.raises(ValueError):
            b2.run(thetas, graph=False)
    finally:
        sp10.configure_sigma(0, 0, -1)
    _close(*b2.run(thetas, graph=True), e0, g0)



def test_batch_all_grouped():
    """A batch whose instances are all large enough to be grouped (no
    persistent launch): (10e,10o) instances in two gate layouts, eagerly and
    over graph replays, equal the per-instance loop to 1e-10."""
    from slowquant_amd import engine as E
    from slowquant_amd.batch import InstanceSpec, build_instances, loop_energy_grad

    specs = [InstanceSpec(10, 5, 5, L, seed) for L, seed in ((1, 31), (2, 32), (1, 33))]
    inst, thetas = build_instances(specs)
    e_l, g_l = loop_energy_grad(inst, thetas)
    b = E.Batch(inst, 4)
    _close(*b.run(thetas), e_l, g_l)
    for _ in range(2):
        _close(*b.run(thetas, graph=True), e_l, g_l)


def test_batch_group_members_own_start_states():
    """Three (10e,10o) instances with one gate layout form one group; each
    member starts from its own random unit state (and one from HF), eagerly
    and over graph replays, equal to Program.energy_grad from the same state
    to 1e-10."""
    import torch

    from slowquant_amd import engine as E
    from slowquant_amd.batch import InstanceSpec, build_instances

    specs = [InstanceSpec(10, 5, 5, 1, seed) for seed in (41, 42, 43)]
    inst, thetas = build_instances(specs)
    b = E.Batch(inst, 1)
    gen = torch.Generator().manual_seed(44)
    starts = []
    for k, (sp, _, _) in enumerate(inst):
        if k == 1:
            starts.append(None)
            continue
        x = torch.randn(sp.dim, generator=gen, dtype=torch.float64)
        starts.append((x / x.norm()).cuda())
    b.set_start_states(starts)
    e_ref, g_ref = [], []
    for (sp, prog, H), th, x in zip(inst, thetas, starts):
        ei, gi = prog.energy_grad(H, th, E.hf_state(sp) if x is None else x)
        e_ref.append(ei)
        g_ref.append(gi)
    _close(*b.run(thetas), e_ref, g_ref)
    for _ in range(2):
        _close(*b.run(thetas, graph=True), e_ref, g_ref)


def test_batch_with_12_orbital_instances():
    """Two (12e,12o) tUPS instances (853776 determinants each, the headline
    space; grouped through the tile-stage and fused row-sigma kernels) beside
    a (6e,6o) instance: eagerly and over graph replays equal to the
    per-instance loop to 1e-10."""
    from slowquant_amd import engine as E
    from slowquant_amd.batch import InstanceSpec, build_instances, loop_energy_grad

    specs = [InstanceSpec(12, 6, 6, 1, 51), InstanceSpec(12, 6, 6, 1, 52), InstanceSpec(6, 3, 3, 2, 53)]
    inst, thetas = build_instances(specs)
    e_l, g_l = loop_energy_grad(inst, thetas)
    b = E.Batch(inst, 4)
    _close(*b.run(thetas), e_l, g_l)
    for _ in range(2):
        _close(*b.run(thetas, graph=True), e_l, g_l)
