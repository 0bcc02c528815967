"""Kernel-shape autotuner: every tile size and row-sigma shape it may choose
gives the default's results (energies 1e-12, gradients and sigma 1e-12
relative), and a tuned choice is written to and read back from the on-disk
cache ($SQ_AUTOTUNE_CACHE)."""

from __future__ import annotations

import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(n, layers, seed):
    from slowquant_amd import engine as E
    from slowquant_amd.synthetic import synthetic_problem
    from slowquant_amd.util import UpsStructure

    lay = UpsStructure()
    lay.create_tiled(n, {"n_layers": layers, "do_tups": True})
    h, g, theta = synthetic_problem(n, 2 * n, lay.n_params, 0.1, seed)
    sp = E.Space(n, n // 2, n // 2)
    prog = E.Program(sp, lay.excitation_operator_type, lay.excitation_indices)
    H = E.Hamiltonian(sp, h, g)
    return sp, prog, H, theta, E.hf_state(sp)


@pytest.mark.parametrize("n", [8, 10])
def test_every_shape_gives_the_default_results(n):
    from slowquant_amd import autotune

    sp, prog, H, theta, psi0 = _problem(n, 2, n)
    e0, g0 = prog.energy_grad(H, theta, psi0, "adjoint")
    psi = prog.state_prep(theta, psi0)
    s0, _ = H.sigma(psi)
    scale = float(s0.abs().max())
    for nt in autotune.TILE_THREADS:
        for t in autotune.TILE_TARGETS:
            prog.set_tile_shape(t, nt)
            e, g = prog.energy_grad(H, theta, psi0, "adjoint")
            assert abs(e - e0) < 1e-12 and np.abs(g - g0).max() < 1e-12, (t, nt)
    prog.set_tile_shape(0, 0)
    for det in (0, 1):
        for r in autotune.SIGMA_ROWS:
            for ns in autotune.SIGMA_SPLITS:
                sp.configure_sigma(r, ns, det)
                s, _ = H.sigma(psi)
                assert float((s - s0).abs().max()) < 1e-12 * scale, (det, r, ns)
    sp.configure_sigma(0, 0, -1)


def test_tune_writes_and_reads_the_cache(tmp_path, monkeypatch):
    from slowquant_amd import autotune

    cache = tmp_path / "tune.json"
    monkeypatch.setenv("SQ_AUTOTUNE_CACHE", str(cache))
    sp, prog, H, theta, psi0 = _problem(8, 1, 3)
    e0, g0 = prog.energy_grad(H, theta, psi0, "adjoint")
    first = autotune.tune(sp, prog, H, theta, psi0, reps=2)
    assert not first["cached"] and first["tile_target"] in autotune.TILE_TARGETS
    stored = json.loads(cache.read_text())
    assert autotune.program_key(sp, prog) in stored
    sp2, prog2, H2, theta2, psi02 = _problem(8, 1, 3)
    second = autotune.tune(sp2, prog2, H2, theta2, psi02, reps=2)
    assert second["cached"] and second["tile_target"] == first["tile_target"]
    e, g = prog2.energy_grad(H2, theta2, psi02, "adjoint")
    assert abs(e - e0) < 1e-12 and np.abs(g - g0).max() < 1e-12
