"""Batched tUPS instances over uneven active spaces.

The reference batches equal-dimension states column-wise (the _SA family of
operator_state_algebra.py, e.g. propagate_state_SA :633-781 and
construct_ups_state_SA :1415-1866); here every instance carries its own
space, Hamiltonian, tUPS layer count and theta, and one persistent launch
(engine.Batch) returns every energy and adjoint gradient.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .synthetic import synthetic_problem
from .util import UpsStructure

# (n_orb, n_alpha, n_beta) of the (4e,4o), (6e,6o), (8e,8o), (10e,10o) spaces
BATCH_SPACES = ((4, 2, 2), (6, 3, 3), (8, 4, 4), (10, 5, 5))


@dataclass
class InstanceSpec:
    n_orb: int
    n_alpha: int
    n_beta: int
    layers: int
    seed: int
    rank: int = 0  # 0: 2 * n_orb
    sigma_theta: float = 0.1


def tups_layout(n_orb: int, layers: int) -> UpsStructure:
    layout = UpsStructure()
    layout.create_tiled(n_orb, {"n_layers": layers, "do_tups": True})
    return layout


def draw_specs(n_inst: int = 512, seed: int = 7, spaces=BATCH_SPACES, layers=(1, 2, 3)) -> list[InstanceSpec]:
    """The 512-instance workload: rng = default_rng(seed); per instance k the
    space (uniform over ``spaces``), then the layer count (uniform over
    ``layers``); instance k's Hamiltonian and theta use seed 1000 + k, rank
    2n, variance 0.3^2/n, theta ~ N(0, 0.1^2)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_inst):
        n, na, nb = spaces[int(rng.integers(len(spaces)))]
        L = int(layers[int(rng.integers(len(layers)))])
        out.append(InstanceSpec(n, na, nb, L, 1000 + k))
    return out


def instance_problem(spec: InstanceSpec):
    """(layout, h, g, theta) of one instance."""
    layout = tups_layout(spec.n_orb, spec.layers)
    rank = spec.rank if spec.rank > 0 else 2 * spec.n_orb
    h, g, theta = synthetic_problem(spec.n_orb, rank, layout.n_params, spec.sigma_theta, spec.seed)
    return layout, h, g, theta


def build_instances(specs):
    """Device objects [(space, program, hamiltonian)] and thetas for ``specs``;
    spaces are shared between instances of equal (n, n_alpha, n_beta)."""
    from . import engine as E

    inst, thetas = [], []
    for s in specs:
        layout, h, g, theta = instance_problem(s)
        sp = E.space_for(s.n_orb, s.n_alpha, s.n_beta)
        prog = E.Program(sp, layout.excitation_operator_type, layout.excitation_indices)
        H = E.Hamiltonian(sp, h, g, 0.0)
        inst.append((sp, prog, H))
        thetas.append(theta)
    return inst, thetas


def loop_energy_grad(instances, thetas):
    """Per-instance reference loop: one adjoint engine call per instance."""
    from . import engine as E

    es, gs = [], []
    for (sp, prog, H), th in zip(instances, thetas):
        e, g = prog.energy_grad(H, th, E.hf_state(sp), "adjoint")
        es.append(e)
        gs.append(g)
    return np.array(es), gs
