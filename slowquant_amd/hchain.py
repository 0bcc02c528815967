"""Minimal molecular front end for hydrogen-only molecules in STO-3G.

Upstream of the state-vector engine: it produces AO integrals and RHF orbitals
for the small H_n workloads (e.g. linear H4 at 1.5 Angstrom spacing).  Only
1s Gaussians occur, so every integral has a closed form (Szabo & Ostlund,
Modern Quantum Chemistry, Appendix A).  The object mimics the attributes of
the reference's SlowQuant front end that IntegralManager reads
(SlowQuant.py:7-49; integral_manager.py:33-125), so it can be passed as
``integral_generator`` to WaveFunctionUPS / WaveFunctionUCC.

Basis data: STO-3G hydrogen (basisset/sto-3g.basis:15-20); Angstrom to bohr
factor 1.889725989 (moleculeclass.py:27).
"""

from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

ANGSTROM_TO_BOHR = 1.889725989
STO3G_H_EXP = np.array([0.3425250914e01, 0.6239137298e00, 0.1688554040e00])
STO3G_H_COEF = np.array([0.1543289673e00, 0.5353281423e00, 0.4446345422e00])


def _boys0(t: np.ndarray) -> np.ndarray:
    t = np.asarray(t, dtype=np.float64)
    out = np.empty_like(t)
    small = t < 1e-12
    out[small] = 1.0 - t[small] / 3.0
    ts = t[~small]
    out[~small] = 0.5 * np.sqrt(np.pi / ts) * erf(np.sqrt(ts))
    return out


class _Molecule:
    def __init__(self, coords: np.ndarray, charge: int) -> None:
        self.coordinates = coords
        self.number_atoms = len(coords)
        self.number_electrons = len(coords) - charge
        z = np.ones(len(coords))
        e = 0.0
        for i in range(len(coords)):
            for j in range(i + 1, len(coords)):
                e += z[i] * z[j] / np.linalg.norm(coords[i] - coords[j])
        self.nuclear_repulsion = e


class _Integrals:
    """Contracted 1s STO-3G integrals over all atoms."""

    def __init__(self, coords: np.ndarray) -> None:
        alpha = STO3G_H_EXP
        norm = (2 * alpha / np.pi) ** 0.75
        d = STO3G_H_COEF * norm
        # normalise the contraction exactly
        s = 0.0
        for a in range(3):
            for b in range(3):
                s += d[a] * d[b] * (np.pi / (alpha[a] + alpha[b])) ** 1.5
        d = d / math.sqrt(s)
        n = len(coords)
        S = np.zeros((n, n))
        T = np.zeros((n, n))
        V = np.zeros((n, n))
        for i in range(n):
            for j in range(n):
                Rab2 = float(np.sum((coords[i] - coords[j]) ** 2))
                for a in range(3):
                    for b in range(3):
                        p = alpha[a] + alpha[b]
                        mu = alpha[a] * alpha[b] / p
                        K = math.exp(-mu * Rab2)
                        cc = d[a] * d[b]
                        s_ab = (np.pi / p) ** 1.5 * K
                        S[i, j] += cc * s_ab
                        T[i, j] += cc * mu * (3 - 2 * mu * Rab2) * s_ab
                        P = (alpha[a] * coords[i] + alpha[b] * coords[j]) / p
                        for c in range(n):
                            Rpc2 = float(np.sum((P - coords[c]) ** 2))
                            V[i, j] += cc * (-2 * np.pi / p) * K * float(_boys0(np.array([p * Rpc2]))[0])
        self.overlap_matrix = S
        self.kinetic_energy_matrix = T
        self.nuclear_attraction_matrix = V
        self.electron_repulsion_tensor = self._eri(coords, alpha, d)
        self._coords = np.asarray(coords, dtype=np.float64)
        self._alpha = alpha
        self._d = d

    def get_multipole_matrix(self, carts: np.ndarray) -> np.ndarray:
        """<a| x^i y^j z^k |b> about the origin for carts = (i, j, k) with
        i + j + k <= 1 (overlap or one dipole component; the front end's
        electric-dipole integrals, integral_manager.py:95-104).  For 1s
        primitives the product a(r) b(r) is a Gaussian centred at
        P = (alpha_a A + alpha_b B) / (alpha_a + alpha_b), so <a|r|b> = s_ab P."""
        carts = np.asarray(carts, dtype=int)
        if carts.shape != (3,) or carts.min() < 0 or carts.sum() > 1:
            raise ValueError("HydrogenSlowQuant provides overlap and dipole integrals only")
        if carts.sum() == 0:
            return self.overlap_matrix.copy()
        k = int(np.argmax(carts))
        coords, alpha, d = self._coords, self._alpha, self._d
        n = len(coords)
        M = np.zeros((n, n))
        for i in range(n):
            for j in range(n):
                Rab2 = float(np.sum((coords[i] - coords[j]) ** 2))
                for a in range(3):
                    for b in range(3):
                        p = alpha[a] + alpha[b]
                        s_ab = (np.pi / p) ** 1.5 * math.exp(-alpha[a] * alpha[b] / p * Rab2)
                        P = (alpha[a] * coords[i, k] + alpha[b] * coords[j, k]) / p
                        M[i, j] += d[a] * d[b] * s_ab * P
        return M

    @staticmethod
    def _eri(coords, alpha, d):
        n = len(coords)
        # primitive-pair quantities, vectorised over (i, a, j, b)
        A = coords[:, None, None, None, :]
        B = coords[None, None, :, None, :]
        aa = alpha[None, :, None, None]
        bb = alpha[None, None, None, :]
        shape = (n, 3, n, 3)
        p = np.broadcast_to(aa + bb, shape)
        Rab2 = np.sum((A - B) ** 2, axis=-1)
        K = np.exp(-aa * bb / p * Rab2)
        P = (aa[..., None] * A + bb[..., None] * B) / p[..., None]
        cc = np.broadcast_to(d[None, :, None, None] * d[None, None, None, :], shape)
        pair_p = p.reshape(-1)
        pair_K = (K * cc).reshape(-1)
        pair_P = P.reshape(-1, 3)
        npair = len(pair_p)
        p1 = pair_p[:, None]
        p2 = pair_p[None, :]
        RPQ2 = np.sum((pair_P[:, None, :] - pair_P[None, :, :]) ** 2, axis=-1)
        val = 2 * np.pi**2.5 / (p1 * p2 * np.sqrt(p1 + p2)) * _boys0(p1 * p2 / (p1 + p2) * RPQ2)
        val = val * pair_K[:, None] * pair_K[None, :]
        val = val.reshape(n, 3, n, 3, n, 3, n, 3).sum(axis=(1, 3, 5, 7))
        assert npair == 9 * n * n
        return val


class _HartreeFock:
    """Restricted Hartree-Fock (same equations as hartreefock_in_memory.py:7-94)."""

    def __init__(self, mol: _Molecule, ints: _Integrals) -> None:
        self.mol_obj = mol
        self.int_obj = ints
        self.de_threshold = 1e-9
        self.rmsd_threshold = 1e-9
        self.max_scf_iterations = 200

    def run_restricted_hartree_fock(self) -> None:
        S = self.int_obj.overlap_matrix
        H = self.int_obj.kinetic_energy_matrix + self.int_obj.nuclear_attraction_matrix
        eri = self.int_obj.electron_repulsion_tensor
        nocc = self.mol_obj.number_electrons // 2
        lam, L = np.linalg.eigh(S)
        X = L @ np.diag(lam**-0.5) @ L.T
        _, Cp = np.linalg.eigh(X.T @ H @ X)
        C = X @ Cp
        D = 2 * C[:, :nocc] @ C[:, :nocc].T
        E_old = float(np.sum(D * H))
        focks, errs = [], []
        for _ in range(self.max_scf_iterations):
            J = np.einsum("pqrs,sr->pq", eri, D)
            K = np.einsum("psqr,sr->pq", eri, D)
            F = H + J - 0.5 * K
            # DIIS on the commutator FDS - SDF
            err = F @ D @ S - S @ D @ F
            focks.append(F)
            errs.append(err)
            focks, errs = focks[-8:], errs[-8:]
            if len(focks) > 1:
                m = len(focks)
                Bm = -np.ones((m + 1, m + 1))
                Bm[m, m] = 0.0
                for i in range(m):
                    for j in range(m):
                        Bm[i, j] = np.sum(errs[i] * errs[j])
                rhs = np.zeros(m + 1)
                rhs[m] = -1.0
                try:
                    c = np.linalg.solve(Bm, rhs)[:m]
                    F = sum(ci * Fi for ci, Fi in zip(c, focks))
                except np.linalg.LinAlgError:
                    pass
            _, Cp = np.linalg.eigh(X.T @ F @ X)
            C = X @ Cp
            Dn = 2 * C[:, :nocc] @ C[:, :nocc].T
            Fn = H + np.einsum("pqrs,sr->pq", eri, Dn) - 0.5 * np.einsum("psqr,sr->pq", eri, Dn)
            E = 0.5 * float(np.sum(Dn * (H + Fn)))
            rmsd = float(np.sqrt(np.sum((Dn - D) ** 2)))
            D = Dn
            done = abs(E - E_old) < self.de_threshold and rmsd < self.rmsd_threshold
            E_old = E
            if done:
                break
        self.E_hf = E_old
        self.mo_coeff = C
        self.rdm1 = D


class HydrogenSlowQuant:
    """SlowQuant-like front end restricted to H atoms and STO-3G."""

    def __init__(self) -> None:
        self.molecule: _Molecule | None = None
        self.integral: _Integrals | None = None
        self.hartree_fock: _HartreeFock | None = None

    def set_molecule(self, molecule_file: str, distance_unit: str = "bohr", molecular_charge: int = 0) -> None:
        unit = {"angstrom": ANGSTROM_TO_BOHR, "bohr": 1.0, "au": 1.0}[distance_unit.lower()]
        coords = []
        for line in molecule_file.replace("\n", ";").split(";"):
            parts = line.split()
            if not parts:
                continue
            if parts[0].lower() != "h":
                raise ValueError("HydrogenSlowQuant supports H atoms only")
            coords.append([float(x) * unit for x in parts[1:4]])
        self.molecule = _Molecule(np.array(coords), molecular_charge)

    def set_basis_set(self, basis: str) -> None:
        if basis.lower() != "sto-3g":
            raise ValueError("HydrogenSlowQuant supports STO-3G only")
        self.integral = _Integrals(self.molecule.coordinates)

    def init_hartree_fock(self) -> None:
        self.hartree_fock = _HartreeFock(self.molecule, self.integral)


def linear_h_chain(n_atoms: int, spacing_angstrom: float) -> HydrogenSlowQuant:
    """Linear H_n along x with RHF orbitals ready."""
    sq = HydrogenSlowQuant()
    sq.set_molecule(";".join(f"H {i * spacing_angstrom} 0.0 0.0" for i in range(n_atoms)), distance_unit="angstrom")
    sq.set_basis_set("sto-3g")
    sq.init_hartree_fock()
    sq.hartree_fock.run_restricted_hartree_fock()
    return sq
