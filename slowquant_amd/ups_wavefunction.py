"""WaveFunctionUPS: drop-in for the reference's unitary product state
(ups_wavefunction.py:38-1194) with the state-vector work on the MI355X engine.

Same constructor, attributes and entry points: thetas (eager state
preparation), ci_coeffs (host numpy, reference convention), energy_elec,
rdm1 / rdm2, _calc_energy_optimization, _calc_gradient_optimization,
run_wf_optimization_1step / 2step, kappa.  The theta gradient uses the
engine's adjoint sweep by default (``gradient_method``: "adjoint", "reference"
for the reference's product-rule sweep, or "plain").
"""

from __future__ import annotations

from collections.abc import Sequence
from typing import Any

import numpy as np

from .density_matrix import get_electronic_energy
from .fermionic_operator import FermionicOperator
from .operators import hamiltonian_0i_0a
from .util import UpsStructure
from .wavefunction_base import WaveFunctionBase


class WaveFunctionUPS(WaveFunctionBase):
    def __init__(
        self,
        cas: Sequence[int],
        mo_coeffs: np.ndarray,
        integral_generator,
        ansatz: str,
        ansatz_options: dict[str, Any] | None = None,
        include_active_kappa: bool = False,
    ) -> None:
        if ansatz_options is None:
            ansatz_options = {}
        self.ansatz_options = ansatz_options
        self._init_spaces(cas, integral_generator, include_active_kappa)
        self.gradient_method = "adjoint"
        self.csf_coeffs = np.zeros(self.num_det)
        hf_det = "1" * self.num_active_elec + "0" * (self.num_active_spin_orbs - self.num_active_elec)
        self._pp = False
        if ansatz.lower() == "tups" and self.ansatz_options.get("do_pp", False):
            pp_det = self._perfect_pairing_det()
            hole = [i for i, (h, p) in enumerate(zip(hf_det, pp_det)) if h == "1" and p == "0"]
            part = [i for i, (h, p) in enumerate(zip(hf_det, pp_det)) if h == "0" and p == "1"]
            hole_sp = sorted({i // 2 + self.num_inactive_orbs for i in hole})
            part_sp = sorted({i // 2 + self.num_inactive_orbs for i in part})
            c = mo_coeffs.copy()
            c[:, hole_sp + part_sp] = c[:, part_sp + hole_sp]
            self._c_mo = c
            self.csf_coeffs[self.ci_info.det2idx[int(pp_det, 2)]] = 1
            self._pp = True
        else:
            self.csf_coeffs[self.ci_info.det2idx[int(hf_det, 2)]] = 1
            self._c_mo = mo_coeffs
        self._ci_host = np.copy(self.csf_coeffs)
        self._psi_dev = None
        self.ups_layout = UpsStructure()
        name = ansatz.lower()
        if name in ("tups", "qnp"):
            self.ansatz_options["do_tups" if name == "tups" else "do_qnp"] = True
            self.ups_layout.create_tiled(self.num_active_orbs, self.ansatz_options)
        elif name in ("fucc", "fuccsd", "ksafupccgsd", "fuccpd", "safuccsd"):
            # The reference accepts "fuccpd"/"safuccsd" here but tests "fuccspd"/
            # "safuccspd" below (ups_wavefunction.py:273-284); kept identical.
            if name == "fuccsd":
                self.ansatz_options["S"] = True
                self.ansatz_options["D"] = True
            elif name == "ksafupccgsd":
                self.ansatz_options["SAGS"] = True
                self.ansatz_options["GpD"] = True
            elif name == "fuccspd":
                self.ansatz_options["pD"] = True
            elif name == "safuccspd":
                self.ansatz_options["SAS"] = True
                self.ansatz_options["SAD"] = True
            self.ansatz_options.setdefault("n_layers", 1)
            self.ups_layout.create_fUCC(
                self.active_occ_idx_shifted, self.active_unocc_idx_shifted, self.active_occ_spin_idx_shifted,
                self.active_unocc_spin_idx_shifted, self.num_active_orbs, self.ansatz_options,
            )
        elif name in ("sdsfuccsd", "ksasdsfupccgsd"):
            self.ansatz_options["D" if name == "sdsfuccsd" else "GpD"] = True
            self.ansatz_options.setdefault("n_layers", 1)
            self.ups_layout.create_SDSfUCC(
                self.active_occ_idx_shifted, self.active_unocc_idx_shifted, self.active_occ_spin_idx_shifted,
                self.active_unocc_spin_idx_shifted, self.num_active_orbs, self.ansatz_options,
            )
        else:
            raise ValueError(f"Got unknown ansatz, {ansatz}")
        self._thetas = np.zeros(self.ups_layout.n_params).tolist()
        self._program = None

    def _perfect_pairing_det(self) -> str:
        """Perfect-pairing reference determinant (ups_wavefunction.py:218-241)."""
        pp = ""
        so = 0
        ne = self.num_active_elec
        n = self.num_active_spin_orbs
        while so < n:
            if ne >= 2 and (n - so) >= 4 and ne <= (n - so - 2):
                pp += "1100"
                ne -= 2
                so += 4
            elif ne == 0:
                pp += "0"
                so += 1
            else:
                pp += "1"
                so += 1
                ne -= 1
        print("perfect-pairing determinant found as:", pp)
        if len(pp) != n or pp.count("1") != self.num_active_elec:
            raise ValueError("Perfect pairing determinant violates orbital or electron numbers")
        return pp

    # ---- device state ------------------------------------------------------
    def _prog(self):
        if self._program is None:
            from . import engine

            self._program = engine.Program(self._space(), self.ups_layout.excitation_operator_type,
                                           self.ups_layout.excitation_indices)
        return self._program

    def _csf_dev(self):
        return self._csf_device()

    def _psi(self):
        """Device state U(theta)|CSF> in the engine convention."""
        if self._psi_dev is None:
            sp = self._space()
            if self._ci_host is not None:
                self._psi_dev = sp.from_reference(self._ci_host)
            else:
                self._psi_dev = self._prog().state_prep(self._thetas, self._csf_dev())
        return self._psi_dev

    @property
    def ci_coeffs(self) -> np.ndarray:
        if self._ci_host is None:
            self._ci_host = self._space().to_reference_numpy(self._psi())
        return self._ci_host

    @ci_coeffs.setter
    def ci_coeffs(self, value) -> None:
        self._ci_host = np.array(value, dtype=np.float64, copy=True)
        self._state_from_thetas = False
        self._psi_dev = None
        self._rdm1 = None
        self._rdm2 = None
        self._energy_elec = None

    @property
    def thetas(self) -> list[float]:
        return self._thetas.copy()

    @thetas.setter
    def thetas(self, theta_vals: list[float]) -> None:
        """ups_wavefunction.py:345-365, which rebuilds the state at once; here
        the state U(theta)|CSF> is prepared on first use (the adjoint gradient
        prepares its own, so an optimiser's gradient call costs no extra
        state preparation).  The thetas already set, with the state still the
        one they give (no ci_coeffs assigned since), change nothing."""
        if len(theta_vals) != len(self._thetas):
            raise ValueError(f"Expected {len(self._thetas)} theta1 values got {len(theta_vals)}")
        if self.__dict__.get("_state_from_thetas", True) and list(theta_vals) == self._thetas:
            return
        self._state_from_thetas = True
        self._rdm1 = None
        self._rdm2 = None
        self._rdm3 = None
        self._rdm4 = None
        self._energy_elec = None
        self._thetas = list(theta_vals).copy()
        self._ci_host = None
        self._psi_dev = None

    # ---- properties computed on the device -----------------------------------
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

    def _get_hamiltonian(self, qiskit_form: bool = False) -> FermionicOperator:
        H = hamiltonian_0i_0a(self.h_mo, self.g_mo, self.num_inactive_orbs, self.num_active_orbs)
        H = H.get_folded_operator(self.num_inactive_orbs, self.num_active_orbs, self.num_virtual_orbs)
        if qiskit_form:
            raise NotImplementedError("qiskit form is outside the state-vector engine")
        return H

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
        elif theta_optimization and self.gradient_method == "adjoint":
            # the optimisers ask for the gradient at the same point next: one
            # adjoint evaluation (state, H psi, reverse sweep) gives both
            H, prog = self._hamiltonian(), self._prog()
            E, g = prog.energy_grad(H, np.asarray(self._thetas, dtype=np.float64), self._csf_dev(), "adjoint")
            self._grad_cache = (np.array(parameters, dtype=np.float64), H, prog, g)
        else:
            _, E = self._hamiltonian().sigma(self._psi())
        self._E_opt_old = E
        self._old_opt_parameters = np.copy(parameters)
        self.num_energy_evals += 1
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
            H, prog = self._hamiltonian(), self._prog()
            if (not kappa_optimization and self.gradient_method == "adjoint" and c is not None and c[1] is H
                    and c[2] is prog and np.array_equal(c[0], np.asarray(parameters, dtype=np.float64))):
                g = c[3]  # from the energy call at these parameters
            else:
                _, g = prog.energy_grad(H, np.asarray(self._thetas, dtype=np.float64), self._csf_dev(),
                                        self.gradient_method)
            gradient[num_kappa:] += g
            self.num_energy_evals += 2 * np.sum(list(self.ups_layout.grad_param_R.values()))
        return gradient

    def energy_and_gradient(self, thetas=None, method: str | None = None) -> tuple[float, np.ndarray]:
        """E(theta) and dE/dtheta in one engine call (no host state copies)."""
        th = self._thetas if thetas is None else list(thetas)
        return self._prog().energy_grad(self._hamiltonian(), np.asarray(th, dtype=np.float64), self._csf_dev(),
                                        method or self.gradient_method)

    def _extra_options(self, optimizer_name: str) -> dict | None:
        if optimizer_name.lower() != "rotosolve":
            return None
        return {
            "R": self.ups_layout.grad_param_R,
            "param_names": self.ups_layout.param_names,
            "f_rotosolve_optimized": self._calc_energy_rotosolve_optimization,
        }

    def _calc_energy_rotosolve_optimization(self, parameters, theta_diffs, theta_idx) -> list[float]:
        """Energies with theta[theta_idx] replaced by each of theta_diffs
        (ups_wavefunction.py:1144-1194, which propagates the shifted states as
        one batch).  All shifted instances run in one batched launch when the
        layout allows it (rotation gates, at most 12 active orbitals: the
        persistent kernel for small spaces, the tile-stage and row-sigma
        kernels over the group of copies from 16384 determinants), else one
        engine energy per shift."""
        th = np.asarray(parameters, dtype=np.float64)
        thetas = np.repeat(th[None, :], len(theta_diffs), axis=0)
        thetas[:, theta_idx] = np.asarray(theta_diffs, dtype=np.float64)
        energies = self._shifted_energies(thetas)
        self.num_energy_evals += len(energies)
        return [float(e) for e in energies]

    def _shifted_batch(self, n: int):
        """Cached batch of n instances of this wave function's program and
        Hamiltonian (started from the perfect-pairing reference when that is
        the CSF), or None where the batched kernel does not apply."""
        from . import engine

        H, prog = self._hamiltonian(), self._prog()
        cache = getattr(self, "_roto_batches", None)
        if cache is None or cache.get("key") != (id(H), id(prog)):
            cache = {"key": (id(H), id(prog)), "H": H}
            self._roto_batches = cache
        if n not in cache:
            try:
                b = engine.Batch([(self._space(), prog, H)] * n)
                if self._pp:
                    b.set_start_states([self._csf_dev()] * n)
                cache[n] = b
            except (ValueError, RuntimeError):
                cache[n] = None
        return cache[n]

    def _shifted_energies(self, thetas: np.ndarray) -> np.ndarray:
        batch = self._shifted_batch(len(thetas))
        if batch is not None:
            e, _ = batch.run(list(thetas))
            return e
        H, prog, csf = self._hamiltonian(), self._prog(), self._csf_dev()
        out = np.zeros(len(thetas))
        for j, t in enumerate(thetas):
            _, out[j] = H.sigma(prog.state_prep(t, csf))
        return out
