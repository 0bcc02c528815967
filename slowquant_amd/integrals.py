"""Integral access and AO -> MO transforms (host side of the engine boundary).

IntegralManager mirrors integral_manager.py:9-125 by duck typing: a
SlowQuant-like object (``molecule.number_electrons``, ``integral.*_matrix``,
``integral.electron_repulsion_tensor``, ``molecule.nuclear_repulsion``) or a
PySCF-like Mole (``nelectron``, ``intor``, ``energy_nuc``).  The transforms
are the einsum contractions of integralfunctions.py:186-211.
"""

from __future__ import annotations

import copy

import numpy as np


def one_electron_integral_transform(C: np.ndarray, int1e: np.ndarray) -> np.ndarray:
    return np.einsum("ai,bj,ab->ij", C, C, int1e, optimize=True)


def two_electron_integral_transform(C: np.ndarray, int2e: np.ndarray) -> np.ndarray:
    return np.einsum("ai,bj,ck,dl,abcd->ijkl", C, C, C, C, int2e, optimize=True)


class IntegralManager:
    __slots__ = ("_dip", "_eri", "_h_ao", "_kin", "_nuc", "int_obj")

    def __init__(self, integral_obj) -> None:
        self.int_obj = copy.deepcopy(integral_obj)
        self._kin = None
        self._nuc = None
        self._eri = None
        self._h_ao = None
        self._dip = None

    def _is_slowquant(self) -> bool:
        return hasattr(self.int_obj, "molecule") and hasattr(self.int_obj, "integral")

    def _is_mole(self) -> bool:
        return hasattr(self.int_obj, "intor") and hasattr(self.int_obj, "nelectron")

    def _unknown(self):
        return ValueError(f"Got unknown integral object, {type(self.int_obj)}")

    @property
    def num_elec(self) -> int:
        if self._is_slowquant():
            return self.int_obj.molecule.number_electrons
        if self._is_mole():
            return self.int_obj.nelectron
        raise self._unknown()

    @property
    def kinetic_energy(self) -> np.ndarray:
        if self._kin is None:
            if self._is_slowquant():
                self._kin = self.int_obj.integral.kinetic_energy_matrix
            elif self._is_mole():
                self._kin = self.int_obj.intor("int1e_kin")
            else:
                raise self._unknown()
        return self._kin

    @property
    def nuclear_electron_attraction(self) -> np.ndarray:
        if self._nuc is None:
            if self._is_slowquant():
                self._nuc = self.int_obj.integral.nuclear_attraction_matrix
            elif self._is_mole():
                self._nuc = self.int_obj.intor("int1e_nuc")
            else:
                raise self._unknown()
        return self._nuc

    @property
    def electron_electron_repulsion(self) -> np.ndarray:
        if self._eri is None:
            if self._is_slowquant():
                self._eri = self.int_obj.integral.electron_repulsion_tensor
            elif self._is_mole():
                self._eri = self.int_obj.intor("int2e")
            else:
                raise self._unknown()
        return self._eri

    @property
    def nuclear_nuclear_repulsion(self) -> float:
        if self._is_slowquant():
            return self.int_obj.molecule.nuclear_repulsion
        if self._is_mole():
            return self.int_obj.energy_nuc()
        raise self._unknown()

    @property
    def electric_dipole(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """AO dipole integrals (x, y, z) (integral_manager.py:94-111)."""
        if self._dip is None:
            if self._is_slowquant():
                self._dip = tuple(self.int_obj.integral.get_multipole_matrix(np.array(c))
                                  for c in ((1, 0, 0), (0, 1, 0), (0, 0, 1)))
            elif self._is_mole():
                x, y, z = self.int_obj.intor("int1e_r", comp=3)
                self._dip = (x, y, z)
            else:
                raise self._unknown()
        return self._dip

    @property
    def h_ao(self) -> np.ndarray:
        if self._h_ao is None:
            self._h_ao = self.nuclear_electron_attraction + self.kinetic_energy
        return self._h_ao
