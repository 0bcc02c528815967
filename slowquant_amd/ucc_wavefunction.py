"""WaveFunctionUCC: drop-in for the reference's unfactorised UCC wave function
(ucc_wavefunction.py:35-1098) with exp(T) applied matrix-free on the GPU.

The reference builds a dense d x d T matrix and calls expm_multiply; here
exp(T(theta))|CSF> is a scaled Taylor series of engine operator applications.
``gradient_method``: "analytic" (default; Gauss-Legendre quadrature of
<e^{-sT} lam|T_i|e^{-sT} psi> over s in [0, 1]) or
"finite_difference" (the reference's forward difference with step
sqrt(eps), ucc_wavefunction.py:1069-1097, evaluated with the same engine).
"""

from __future__ import annotations

from collections.abc import Sequence

import numpy as np

from .density_matrix import get_electronic_energy
from .util import UccStructure
from .wavefunction_base import WaveFunctionBase


class WaveFunctionUCC(WaveFunctionBase):
    def __init__(
        self,
        cas: Sequence[int],
        mo_coeffs: np.ndarray,
        integral_generator,
        excitations: str,
        include_active_kappa: bool = False,
    ) -> None:
        self._init_spaces(cas, integral_generator, include_active_kappa)
        self._c_mo = mo_coeffs
        self.csf_coeffs = np.zeros(self.num_det)
        hf_det = int("1" * self.num_active_elec + "0" * (self.num_active_spin_orbs - self.num_active_elec), 2)
        self.csf_coeffs[self.ci_info.det2idx[hf_det]] = 1
        self._ci_coeffs = np.copy(self.csf_coeffs)
        self._psi_dev = None
        self._excitations = excitations
        self.ucc_layout = UccStructure()
        ex = excitations.lower()
        if "s" in ex:
            self.ucc_layout.add_sa_singles(self.active_occ_idx_shifted, self.active_unocc_idx_shifted)
        if "d" in ex:
            self.ucc_layout.add_sa_doubles(self.active_occ_idx_shifted, self.active_unocc_idx_shifted)
        if "t" in ex:
            self.ucc_layout.add_triples(self.active_occ_spin_idx_shifted, self.active_unocc_spin_idx_shifted)
        if "q" in ex:
            self.ucc_layout.add_quadruples(self.active_occ_spin_idx_shifted, self.active_unocc_spin_idx_shifted)
        if "5" in ex:
            self.ucc_layout.add_quintuples(self.active_occ_spin_idx_shifted, self.active_unocc_spin_idx_shifted)
        if "6" in ex:
            self.ucc_layout.add_sextuples(self.active_occ_spin_idx_shifted, self.active_unocc_spin_idx_shifted)
        self._thetas = np.zeros(self.ucc_layout.n_params).tolist()
        self._program = None
        self.gradient_method = "analytic"

    # ---- device -------------------------------------------------------------
    def _prog(self):
        if self._program is None:
            from . import engine

            types = ["sa_single_ucc" if t == "sa_single" else t for t in self.ucc_layout.excitation_operator_type]
            self._program = engine.Program(self._space(), types, self.ucc_layout.excitation_indices)
        return self._program

    def _expm(self, theta, psi_in, t: float = 1.0):
        from . import engine as E

        return E.ucc_expm(self._prog(), theta, psi_in, t)

    def _psi(self):
        if self._psi_dev is None:
            self._psi_dev = self._expm(np.asarray(self._thetas, dtype=np.float64), self._csf_device())
        return self._psi_dev

    @property
    def ci_coeffs(self) -> np.ndarray:
        if self._ci_coeffs is None:
            self._ci_coeffs = self._space().to_reference_numpy(self._psi())
        return self._ci_coeffs

    @property
    def thetas(self) -> list[float]:
        return self._thetas.copy()

    @thetas.setter
    def thetas(self, theta: list[float]) -> None:
        """ucc_wavefunction.py:278-293; the thetas already set change nothing
        (the state and the quantities built from it stay valid)."""
        if len(theta) != len(self._thetas):
            raise ValueError(f"Expected {len(self._thetas)} theta1 values got {len(theta)}")
        if list(theta) == self._thetas:
            return
        self._rdm1 = None
        self._rdm2 = None
        self._rdm3 = None
        self._rdm4 = None
        self._energy_elec = None
        self._ci_coeffs = None
        self._psi_dev = None
        self._thetas = list(theta).copy()

    @property
    def rdm1(self) -> np.ndarray:
        if self._rdm1 is None:
            from . import engine

            self._rdm1, self._rdm2 = engine.rdm12(self._space(), self._psi())
        return self._rdm1

    @property
    def rdm2(self) -> np.ndarray:
        if self._rdm2 is None:
            from . import engine

            self._rdm1, self._rdm2 = engine.rdm12(self._space(), self._psi())
        return self._rdm2

    @property
    def energy_elec(self) -> float:
        if self._energy_elec is None:
            _, self._energy_elec = self._hamiltonian().sigma(self._psi())
        return self._energy_elec

    def energy_and_gradient(self, thetas=None):
        """Energy and analytic dE/dtheta in one engine call."""
        from . import engine as E

        th = self._thetas if thetas is None else thetas
        return E.ucc_energy_grad(self._prog(), self._hamiltonian(), th, self._csf_device())

    def _finite_difference_gradient(self) -> np.ndarray:
        """The reference's forward difference (ucc_wavefunction.py:1069-1097):
        g_i = 2 (<exp(T(theta) + h T_i) csf | H psi> - E) / h, h = sqrt(eps)."""
        psi = self._psi()
        lam, E = self._hamiltonian().sigma(psi)
        eps = np.finfo(np.float64).eps ** 0.5
        csf = self._csf_device()
        th = np.asarray(self._thetas, dtype=np.float64)
        g = np.zeros(len(th))
        for i in range(len(th)):
            tp = th.copy()
            tp[i] += eps  # T(theta) + T(eps e_i) = T(theta + eps e_i): T is linear in theta
            bra = self._expm(tp, csf)
            g[i] = 2.0 * (float((bra * lam).sum()) - E) / eps
        return g

    # ---- optimisation entry points -------------------------------------------
    def _calc_energy_optimization(self, parameters, theta_optimization: bool, kappa_optimization: bool) -> float:
        if np.max(np.abs(np.array(self._old_opt_parameters) - np.array(parameters))) < 10**-14:
            return self._E_opt_old
        num_kappa = 0
        if kappa_optimization:
            num_kappa = len(self.kappa_idx)
            self.kappa = list(parameters[:num_kappa])
        if theta_optimization:
            self.thetas = list(parameters[num_kappa:])
        if kappa_optimization:
            E = get_electronic_energy(self.h_mo, self.g_mo, self.num_inactive_orbs, self.num_active_orbs,
                                      self.rdm1, self.rdm2)
        elif theta_optimization and self.gradient_method == "analytic":
            # the optimisers ask for the gradient at the same point next: one
            # engine call gives the energy and the analytic gradient
            H, prog = self._hamiltonian(), self._prog()
            E, g = self.energy_and_gradient()
            self._grad_cache = (np.array(parameters, dtype=np.float64), H, prog, g)
        else:
            E = self.energy_elec
        self._E_opt_old = E
        self._old_opt_parameters = np.copy(parameters)
        return E

    def _calc_gradient_optimization(self, parameters, theta_optimization: bool, kappa_optimization: bool) -> np.ndarray:
        gradient = np.zeros(len(parameters))
        num_kappa = 0
        if kappa_optimization:
            num_kappa = len(self.kappa_idx)
            self.kappa = list(parameters[:num_kappa])
        if theta_optimization:
            self.thetas = list(parameters[num_kappa:])
        if kappa_optimization:
            gradient[:num_kappa] = self._orbital_gradient(self.rdm1, self.rdm2)
        if theta_optimization:
            c = getattr(self, "_grad_cache", None)
            if self.gradient_method == "finite_difference":
                gradient[num_kappa:] = self._finite_difference_gradient()
            elif (not kappa_optimization and c is not None and c[1] is self._hamiltonian() and c[2] is self._prog()
                  and np.array_equal(c[0], np.asarray(parameters, dtype=np.float64))):
                gradient[num_kappa:] = c[3]  # from the energy call at these parameters
            else:
                gradient[num_kappa:] = self.energy_and_gradient()[1]
        return gradient

    def _reset_kappa_after_two_step(self) -> None:
        # The reference writes the zeros into a copy of kappa
        # (ucc_wavefunction.py:836-838); only the expansion point is reset.
        for i in range(len(self.kappa)):
            self._kappa_old[i] = 0.0
