"""WaveFunctionSAUPS: drop-in for the reference's state-averaged unitary
product state (sa_ups_wavefunction.py:39-1082) on the MI355X engine.

One U(theta) acts on K orthonormal reference states |CSF_k>; the averaged
energy (1/K) sum_k <CSF_k|U^T H U|CSF_k> is minimised and the states are then
mixed by diagonalising H in the span of U|CSF_k> (subspace diagonalisation,
sa_ups_wavefunction.py:741-769).  Every state-vector operation runs on the
device: each state is prepared by the engine's fused gate stages, H psi_k by
the FP64 MFMA sigma, the averaged theta gradient by the adjoint sweep per
state, and the averaged RDMs by the MFMA Gram kernel per state.

Same constructor, attributes and entry points as the reference: states,
ci_coeffs (host array, K x dim, reference convention), thetas (lazy state
preparation, as in the reference), rdm1 / rdm2 (state averaged), sa_energy,
energy_states, excitation_energies, get_transition_property,
get_oscillator_strenghts (the reference's spelling), _calc_energy_optimization
(with return_all_states), _calc_gradient_optimization,
_calc_energy_rotosolve_optimization, run_wf_optimization_1step / 2step.
"""

from __future__ import annotations

from collections.abc import Sequence
from typing import Any

import numpy as np
import scipy.linalg

from .util import UpsStructure
from .wavefunction_base import WaveFunctionBase



# Largest CI dimension for which the K states' energies and gradients run as
# one batched launch (one workgroup per state) rather than state by state on
# the whole GPU.
_BATCH_MAX_DIM = 20000

class WaveFunctionSAUPS(WaveFunctionBase):
    def __init__(
        self,
        cas: Sequence[int],
        mo_coeffs: np.ndarray,
        integral_generator,
        states: tuple[list[list[float]], list[list[str]]],
        ansatz: str,
        ansatz_options: dict[str, Any] | None = None,
        include_active_kappa: bool = False,
    ) -> None:
        """Constructor checks and layouts of sa_ups_wavefunction.py:40-289."""
        if ansatz_options is None:
            ansatz_options = {}
        self.ansatz_options = ansatz_options
        self._init_spaces(cas, integral_generator, include_active_kappa)
        self._c_mo = mo_coeffs
        self._states = states
        self.gradient_method = "adjoint"
        self._sa_energy = None
        self._state_energies = None
        self._state_ci_coeffs = None
        self.num_states = len(states[0])
        self.csf_coeffs = np.zeros((self.num_states, self.num_det))
        for i, (coeffs, on_vecs) in enumerate(zip(states[0], states[1])):
            if len(coeffs) != len(on_vecs):
                raise ValueError(
                    f"Mismatch in number of coefficients, {len(coeffs)}, and number of determinants, "
                    f"{len(on_vecs)}. For {coeffs} and {on_vecs}"
                )
            for coeff, on_vec in zip(coeffs, on_vecs):
                if len(on_vec) != self.num_active_spin_orbs:
                    raise ValueError(
                        f"Length of determinant, {len(on_vec)}, does not match number of active spin orbitals, "
                        f"{self.num_active_spin_orbs}. For determinant, {on_vec}"
                    )
                self.csf_coeffs[i, self.ci_info.det2idx[int(on_vec, 2)]] = coeff
        overlap = self.csf_coeffs @ self.csf_coeffs.T
        for i in range(self.num_states):
            for j in range(self.num_states):
                if i == j and abs(1 - overlap[i, i]) > 10**-10:
                    raise ValueError(f"state {i} is not normalized got overlap of {overlap[i, i]}")
                if i != j and abs(overlap[i, j]) > 10**-10:
                    raise ValueError(f"state {i} and {j} are not orthogonal got overlap of {overlap[i, j]}")
        self._ci_host = np.copy(self.csf_coeffs)
        self._psis_dev = None
        self.ups_layout = UpsStructure()
        name = ansatz.lower()
        if name in ("tups", "qnp"):
            if self.ansatz_options.get("do_pp"):
                raise ValueError("perfect pairing is not supported for Ansatz in SA UPS wave functions.")
            self.ansatz_options["do_tups" if name == "tups" else "do_qnp"] = True
            self.ups_layout.create_tiled(self.num_active_orbs, self.ansatz_options)
        elif name in ("fucc", "ksafupccgsd", "safuccsd"):
            if name == "ksafupccgsd":
                self.ansatz_options["SAGS"] = True
                self.ansatz_options["GpD"] = True
            elif name == "safuccsd":
                self.ansatz_options["SAS"] = True
                self.ansatz_options["SAD"] = True
            self.ansatz_options.setdefault("n_layers", 1)
            self.ups_layout.create_fUCC(
                self.active_occ_idx_shifted, self.active_unocc_idx_shifted, self.active_occ_spin_idx_shifted,
                self.active_unocc_spin_idx_shifted, self.num_active_orbs, self.ansatz_options,
            )
        elif name == "ksasdsfupccgsd":
            self.ansatz_options["GpD"] = True
            self.ansatz_options.setdefault("n_layers", 1)
            self.ups_layout.create_SDSfUCC(
                self.active_occ_idx_shifted, self.active_unocc_idx_shifted, self.active_occ_spin_idx_shifted,
                self.active_unocc_spin_idx_shifted, self.num_active_orbs, self.ansatz_options,
            )
        else:
            raise ValueError(f"Got unknown ansatz, {ansatz}")
        self._thetas = np.zeros(self.ups_layout.n_params).tolist()
        self._program = None
        self._csf_dev_cache = None

    # ---- device state ------------------------------------------------------
    def _prog(self):
        if self._program is None:
            from . import engine

            self._program = engine.Program(self._space(), self.ups_layout.excitation_operator_type,
                                           self.ups_layout.excitation_indices)
        return self._program

    def _csfs_dev(self):
        """The K reference states on the device (engine convention)."""
        if self._csf_dev_cache is None:
            sp = self._space()
            self._csf_dev_cache = [sp.from_reference(c) for c in self.csf_coeffs]
        return self._csf_dev_cache

    def _psis(self):
        """U(theta)|CSF_k> for every state, on the device."""
        if self._psis_dev is None:
            sp = self._space()
            if self._ci_host is not None:
                self._psis_dev = [sp.from_reference(c) for c in self._ci_host]
            else:
                th = np.asarray(self._thetas, dtype=np.float64)
                self._psis_dev = [self._prog().state_prep(th, c) for c in self._csfs_dev()]
        return self._psis_dev

    def _reset_state(self) -> None:
        self._rdm1 = None
        self._rdm2 = None
        self._sa_energy = None
        self._state_energies = None
        self._state_ci_coeffs = None

    # ---- reference attributes ----------------------------------------------
    @property
    def kappa(self) -> list[float]:
        return self._kappa.copy()

    @kappa.setter
    def kappa(self, k: list[float]) -> None:
        """Moves the expansion point (sa_ups_wavefunction.py:296-311); the
        kappa already set, with no expansion point pending, changes nothing."""
        if list(k) == self._kappa and self._kappa_old == self._kappa:
            return
        self._h_mo = None
        self._g_mo = None
        self.__dict__["_mo_dev_cache"] = {}
        self._dev_ham = None  # built from the integrals just dropped
        self._sa_energy = None
        self._state_energies = None
        self._kappa = list(k).copy()
        self._c_mo = self.c_mo
        self._kappa_old = self.kappa
        self._state_ci_coeffs = None

    @property
    def ci_coeffs(self) -> np.ndarray:
        """States U(theta)|CSF_k>, K x dim host array (sa_ups_wavefunction.py:313-327)."""
        if self._ci_host is None:
            sp = self._space()
            self._ci_host = np.stack([sp.to_reference_numpy(p) for p in self._psis()])
        return self._ci_host

    @property
    def thetas(self) -> list[float]:
        return self._thetas.copy()

    @thetas.setter
    def thetas(self, theta_vals: list[float]) -> None:
        """sa_ups_wavefunction.py:338-353; the states are prepared on first use."""
        if len(theta_vals) != len(self._thetas):
            raise ValueError(f"Expected {len(self._thetas)} theta1 values got {len(theta_vals)}")
        self._reset_state()
        self._thetas = list(theta_vals).copy()
        self._ci_host = None
        self._psis_dev = None

    @property
    def rdm1(self) -> np.ndarray:
        """State-averaged active-space 1-RDM (sa_ups_wavefunction.py:397-424)."""
        if self._rdm1 is None:
            self._state_rdms()
        return self._rdm1

    @property
    def rdm2(self) -> np.ndarray:
        """State-averaged active-space 2-RDM (sa_ups_wavefunction.py:426-479)."""
        if self._rdm2 is None:
            self._state_rdms()
        return self._rdm2

    def _state_rdms(self) -> None:
        from . import engine

        n = self.num_active_orbs
        r1 = np.zeros((n, n))
        r2 = np.zeros((n, n, n, n))
        for psi in self._psis():
            a, b = engine.rdm12(self._space(), psi)
            r1 += a
            r2 += b
        self._rdm1 = r1 / self.num_states
        self._rdm2 = r2 / self.num_states

    def _state_batch(self):
        """The K states' energies and adjoint gradients as one persistent launch
        (instances started from each reference state), where the batched
        kernel applies and pays: rotation gates, at most _BATCH_MAX_DIM
        determinants (one workgroup per state; larger spaces run each state
        across the whole GPU instead).  None otherwise."""
        if self._space().dim > _BATCH_MAX_DIM:
            return None
        return self._shifted_batch(1)

    def _batch_run(self):
        """(energies[K], gradients[K]) of the current theta from the state batch."""
        th = np.asarray(self._thetas, dtype=np.float64)
        H = self._hamiltonian()
        hit = getattr(self, "_batch_result", None)
        if hit is not None and hit[0] == th.tobytes() and hit[1] is H:
            return hit[2]
        res = self._state_batch().run([th] * self.num_states, graph=True)
        self._batch_result = (th.tobytes(), H, res)  # H kept: compared by identity
        return res

    def _loop_run(self):
        """(energies[K], gradients[K]) of the current theta, one adjoint
        evaluation per state (spaces without the state batch), kept for the
        optimiser's gradient call at the same theta."""
        th = np.asarray(self._thetas, dtype=np.float64)
        H, prog = self._hamiltonian(), self._prog()
        hit = getattr(self, "_loop_result", None)
        if hit is not None and hit[0] == th.tobytes() and hit[1] is H and hit[2] is prog:
            return hit[3]
        out = [prog.energy_grad(H, th, csf, "adjoint") for csf in self._csfs_dev()]
        res = (np.array([e for e, _ in out]), [g for _, g in out])
        self._loop_result = (th.tobytes(), H, prog, res)  # H and prog kept: compared by identity
        return res

    def _state_energy_list(self) -> np.ndarray:
        if self._state_batch() is not None:
            return np.asarray(self._batch_run()[0], dtype=np.float64)
        H = self._hamiltonian()
        return np.array([H.sigma(p)[1] for p in self._psis()])

    @property
    def sa_energy(self) -> float:
        """(1/K) sum_k <psi_k|H|psi_k> (sa_ups_wavefunction.py:495-515)."""
        if self._sa_energy is None:
            self._sa_energy = float(np.mean(self._state_energy_list()))
        return self._sa_energy

    def _subspace_matrix(self, H) -> np.ndarray:
        """M_ij = <psi_i|O|psi_j> for a device operator O (sigma form)."""
        sp = self._space()
        bras = self.ci_coeffs
        M = np.zeros((self.num_states, self.num_states))
        for j, psi in enumerate(self._psis()):
            col = sp.to_reference_numpy(H.sigma(psi, want_energy=False)[0])
            M[:, j] = bras @ col
        return M

    def _do_state_ci(self) -> None:
        """Diagonalise H in span{psi_k} (sa_ups_wavefunction.py:741-769).  The
        subspace matrix is symmetrised and diagonalised with eigh: the
        eigenvalues are those of the reference's eig + argsort; eigenvectors
        agree up to sign (and rotation inside degenerate pairs), which leaves
        energies and oscillator strengths unchanged."""
        M = self._subspace_matrix(self._hamiltonian())
        M = 0.5 * (M + M.T)
        eigval, eigvec = scipy.linalg.eigh(M)
        self._state_energies = eigval
        self._state_ci_coeffs = eigvec

    @property
    def energy_states(self) -> np.ndarray:
        if self._state_energies is None:
            self._do_state_ci()
        return self._state_energies

    @property
    def excitation_energies(self) -> np.ndarray:
        e = self.energy_states
        return np.array(e[1:] - e[0])

    def get_transition_property(self, ao_integral: np.ndarray) -> np.ndarray:
        """t_n = <0|O|n> between the subspace-diagonalised states for a
        one-electron operator O (sa_ups_wavefunction.py:797-833).  O folded
        into the active space is 2 sum_i o_ii (inactive) + sum_tu o_tu E_tu,
        applied on the device as a one-body operator with that constant."""
        from . import engine

        if self._state_ci_coeffs is None:
            self._do_state_ci()
        from . import orbitals

        o = orbitals.ao2mo(self.c_mo, ao_integral, None)[0].cpu().numpy()
        ni, na = self.num_inactive_orbs, self.num_active_orbs
        o_act = np.array(o[ni:ni + na, ni:ni + na])
        o_act[np.abs(o_act) <= 1e-14] = 0.0
        const = sum(2.0 * o[i, i] for i in range(ni) if abs(o[i, i]) > 1e-14)
        O = engine.Hamiltonian(self._space(), o_act, np.zeros((na, na, na, na)), const)
        M = self._subspace_matrix(O)
        C = self._state_ci_coeffs
        return np.array([C[:, i + 1] @ M @ C[:, 0] for i in range(self.num_states - 1)])

    def get_oscillator_strenghts(self) -> np.ndarray:
        """f_n = 2/3 eps_n |<0|mu|n>|^2 (sa_ups_wavefunction.py:835-853)."""
        dx, dy, dz = self.int_gen.electric_dipole
        tx = self.get_transition_property(dx)
        ty = self.get_transition_property(dy)
        tz = self.get_transition_property(dz)
        return 2.0 / 3.0 * self.excitation_energies * (tx**2 + ty**2 + tz**2)

    # ---- optimisation entry points -------------------------------------------
    def _calc_energy_optimization(self, parameters, theta_optimization: bool, kappa_optimization: bool,
                                  return_all_states: bool = False):
        """sa_ups_wavefunction.py:855-914."""
        if np.max(np.abs(np.array(self._old_opt_parameters) - np.array(parameters))) < 10**-14:
            return self._E_opt_old
        num_kappa = 0
        if kappa_optimization:
            num_kappa = len(self.kappa_idx)
            self.kappa = list(parameters[:num_kappa])
        if theta_optimization:
            self.thetas = list(parameters[num_kappa:])
        if (theta_optimization and not kappa_optimization and self.gradient_method == "adjoint"
                and self._state_batch() is None):
            energies = self._loop_run()[0]  # the gradient call at these parameters reuses the evaluations
        else:
            energies = self._state_energy_list()
        if return_all_states:
            self._E_opt_old = np.copy(energies)
            self._old_opt_parameters = np.copy(parameters)
            return energies
        E = float(np.mean(energies))
        self._E_opt_old = E
        self._old_opt_parameters = np.copy(parameters)
        self.num_energy_evals += self.num_states
        return E

    def _calc_gradient_optimization(self, parameters, theta_optimization: bool, kappa_optimization: bool) -> np.ndarray:
        """State-averaged (kappa, theta) gradient (sa_ups_wavefunction.py:916-1003):
        the theta part is the mean of the per-state adjoint gradients."""
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
            H = self._hamiltonian()
            th = np.asarray(self._thetas, dtype=np.float64)
            g = np.zeros(len(th))
            if self.gradient_method == "adjoint" and self._state_batch() is not None:
                for gk in self._batch_run()[1]:
                    g += gk
            elif self.gradient_method == "adjoint":
                for gk in self._loop_run()[1]:
                    g += gk
            else:
                for csf in self._csfs_dev():
                    g += self._prog().energy_grad(H, th, csf, self.gradient_method)[1]
            gradient[num_kappa:] += g / self.num_states
            self.num_energy_evals += 2 * np.sum(list(self.ups_layout.grad_param_R.values())) * self.num_states
        return gradient

    def _extra_options(self, optimizer_name: str) -> dict | None:
        if optimizer_name.lower() != "rotosolve":
            return None
        return {
            "R": self.ups_layout.grad_param_R,
            "param_names": self.ups_layout.param_names,
            "f_rotosolve_optimized": self._calc_energy_rotosolve_optimization,
        }

    def _calc_energy_rotosolve_optimization(self, parameters, theta_diffs, theta_idx) -> list[float]:
        """For each value of theta[theta_idx] in theta_diffs: the SUM over the
        states of <psi_k|H|psi_k> (sa_ups_wavefunction.py:1005-1082 sums and
        does not divide by K; RotoSolve only reconstructs the curve's shape,
        so the scale does not move its minimum)."""
        th = np.asarray(parameters, dtype=np.float64)
        n, K = len(theta_diffs), self.num_states
        thetas = np.repeat(th[None, :], n, axis=0)
        thetas[:, theta_idx] = np.asarray(theta_diffs, dtype=np.float64)
        batch = self._shifted_batch(n)
        if batch is not None:
            e, _ = batch.run(list(np.repeat(thetas, K, axis=0)))
            out = [float(v) for v in e.reshape(n, K).sum(axis=1)]
        else:
            H, prog, csfs = self._hamiltonian(), self._prog(), self._csfs_dev()
            out = [float(sum(H.sigma(prog.state_prep(t, c))[1] for c in csfs)) for t in thetas]
        self.num_energy_evals += self.num_states
        return out

    def _shifted_batch(self, n: int):
        """Cached batch of n x K instances (shift-major, state-minor), each
        started from its reference state, or None where the batched kernel
        does not apply (non-rotation gates, more than 12 orbitals)."""
        from . import engine

        H, prog = self._hamiltonian(), self._prog()
        cache = getattr(self, "_roto_batches", None)
        if cache is None or cache.get("key") != (id(H), id(prog)):
            cache = {"key": (id(H), id(prog)), "H": H}
            self._roto_batches = cache
        if n not in cache:
            try:
                b = engine.Batch([(self._space(), prog, H)] * (n * self.num_states))
                b.set_start_states(self._csfs_dev() * n)
                cache[n] = b
            except (ValueError, RuntimeError):
                cache[n] = None
        return cache[n]

    def run_wf_optimization_2step(self, optimizer_name: str, orbital_optimization: bool = False, tol: float = 1e-10,
                                  maxiter: int = 1000, is_silent_subiterations: bool = False) -> None:
        """sa_ups_wavefunction.py:517-638: the shared two-step driver, then the
        subspace diagonalisation."""
        super().run_wf_optimization_2step(optimizer_name, orbital_optimization, tol, maxiter, is_silent_subiterations)
        self._after_optimization()

    def run_wf_optimization_1step(self, optimizer_name: str, orbital_optimization: bool = False, tol: float = 1e-10,
                                  maxiter: int = 1000) -> None:
        """sa_ups_wavefunction.py:640-739."""
        super().run_wf_optimization_1step(optimizer_name, orbital_optimization, tol, maxiter)
        self._after_optimization()

    def _after_optimization(self) -> None:
        e = self._energy_elec
        self._do_state_ci()
        self._sa_energy = e
