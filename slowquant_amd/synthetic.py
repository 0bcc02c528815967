"""Seeded synthetic active-space Hamiltonians and a duck-typed integral generator.

Draw order (one numpy Generator per instance): h, then the R symmetric
factors B^L, then theta.
  h = triu(A) + triu(A,1)^T with A ~ N(0, 0.5^2), h_pp -= (2.0 - 0.25 p)
  B^L = triu(M) + triu(M,1)^T with M ~ N(0, (0.3/sqrt(n))^2)
  g_pqrs = sum_L B^L_pq B^L_rs   (8-fold symmetric, positive semidefinite)
  theta ~ N(0, sigma_theta^2)
No inactive or virtual orbitals; with identity MO coefficients h_mo = h and
g_mo = g exactly.
"""

from __future__ import annotations

import numpy as np


def synthetic_hamiltonian(n: int, rank: int, rng: np.random.Generator) -> tuple[np.ndarray, np.ndarray]:
    A = rng.normal(0.0, 0.5, (n, n))
    h = np.triu(A) + np.triu(A, 1).T
    h[np.diag_indices(n)] += -(2.0 - 0.25 * np.arange(n))
    B = np.empty((rank, n, n))
    for L in range(rank):
        M = rng.normal(0.0, 0.3 / np.sqrt(n), (n, n))
        B[L] = np.triu(M) + np.triu(M, 1).T
    g = np.einsum("Lpq,Lrs->pqrs", B, B)
    return h, g


def synthetic_problem(n: int, rank: int, n_theta: int, sigma_theta: float, seed: int):
    """(h, g, theta) for one instance, in the documented draw order."""
    rng = np.random.default_rng(seed)
    h, g = synthetic_hamiltonian(n, rank, rng)
    theta = rng.normal(0.0, sigma_theta, n_theta)
    return h, g, theta


class _Mol:
    def __init__(self, n_elec: int, e_nuc: float) -> None:
        self.number_electrons = n_elec
        self.nuclear_repulsion = e_nuc


class _Ints:
    def __init__(self, h: np.ndarray, g: np.ndarray) -> None:
        n = h.shape[0]
        self.overlap_matrix = np.eye(n)
        self.kinetic_energy_matrix = np.zeros((n, n))
        self.nuclear_attraction_matrix = np.array(h, dtype=np.float64)
        self.electron_repulsion_tensor = np.array(g, dtype=np.float64)


class SyntheticIntegrals:
    """Integral generator exposing the SlowQuant attributes IntegralManager reads
    (integral_manager.py:33-125): kinetic = 0, nuclear attraction = h, eri = g,
    so that with mo_coeffs = I the MO integrals are h and g exactly."""

    def __init__(self, h: np.ndarray, g: np.ndarray, n_elec: int, e_nuc: float = 0.0) -> None:
        self.molecule = _Mol(n_elec, e_nuc)
        self.integral = _Ints(h, g)
