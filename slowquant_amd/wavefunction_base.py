"""Shared orbital-space bookkeeping of the UPS and UCC wave functions.

Follows the constructor logic of ups_wavefunction.py:39-205 /
ucc_wavefunction.py:36-200: active space chosen around the Fermi level,
inactive / virtual spin-orbital lists, shifted active indices, non-redundant
orbital-rotation (kappa) index lists, and the moving expansion point of the
orbital rotations (c_mo = C_old expm(-kappa)).  Device-side pieces (CI space,
folded Hamiltonian, RDMs) are created lazily so construction needs no GPU.
"""

from __future__ import annotations

from collections.abc import Sequence

import numpy as np

import time
from functools import partial

from .ci_spaces import get_indexing
from .optimizers import Optimizers
from .integrals import IntegralManager


class WaveFunctionBase:
    def _init_spaces(self, cas: Sequence[int], integral_generator, include_active_kappa: bool) -> None:
        if len(cas) != 2:
            raise ValueError(f"cas must have two elements, got {len(cas)} elements.")
        self.int_gen = IntegralManager(integral_generator)
        self.num_orbs = len(self.int_gen.kinetic_energy)
        self.num_spin_orbs = 2 * self.num_orbs
        self._include_active_kappa = include_active_kappa
        self._rdm1 = None
        self._rdm2 = None
        self._h_mo = None
        self._g_mo = None
        self._energy_elec = None
        self.num_energy_evals = 0
        n_el = self.int_gen.num_elec
        # active spin orbitals: cas[0] electrons below the Fermi level, the rest above
        active = list(range(n_el - cas[0], n_el)) + list(range(n_el, n_el + 2 * cas[1] - cas[0]))
        act_set = set(active)
        self.inactive_spin_idx = [i for i in range(n_el) if i not in act_set]
        self.active_occ_spin_idx = [i for i in range(n_el) if i in act_set]
        self.active_unocc_spin_idx = [i for i in range(n_el, self.num_spin_orbs) if i in act_set]
        self.virtual_spin_idx = [i for i in range(n_el, self.num_spin_orbs) if i not in act_set]
        self.active_spin_idx = self.active_occ_spin_idx + self.active_unocc_spin_idx
        self.num_inactive_spin_orbs = len(self.inactive_spin_idx)
        self.num_active_spin_orbs = len(self.active_spin_idx)
        self.num_virtual_spin_orbs = len(self.virtual_spin_idx)
        self.num_active_elec = len(self.active_occ_spin_idx)
        if self.num_active_elec % 2 != 0:
            raise ValueError("Number of active electrons has to be even")
        self.num_active_elec_alpha = self.num_active_elec // 2
        self.num_active_elec_beta = self.num_active_elec // 2
        self.num_inactive_orbs = self.num_inactive_spin_orbs // 2
        self.num_active_orbs = self.num_active_spin_orbs // 2
        self.num_virtual_orbs = self.num_virtual_spin_orbs // 2

        def spatial(lst):
            out = []
            for i in lst:
                if i // 2 not in out:
                    out.append(i // 2)
            return out

        self.inactive_idx = spatial(self.inactive_spin_idx)
        self.active_idx = spatial(self.active_spin_idx)
        self.virtual_idx = spatial(self.virtual_spin_idx)
        self.active_occ_idx = spatial(self.active_occ_spin_idx)
        self.active_unocc_idx = spatial(self.active_unocc_spin_idx)
        s0 = min(self.active_spin_idx) if self.active_spin_idx else 0
        self.active_spin_idx_shifted = [i - s0 for i in self.active_spin_idx]
        self.active_occ_spin_idx_shifted = [i - s0 for i in self.active_occ_spin_idx]
        self.active_unocc_spin_idx_shifted = [i - s0 for i in self.active_unocc_spin_idx]
        p0 = min(self.active_idx) if self.active_idx else 0
        self.active_idx_shifted = [i - p0 for i in self.active_idx]
        self.active_occ_idx_shifted = [i - p0 for i in self.active_occ_idx]
        self.active_unocc_idx_shifted = [i - p0 for i in self.active_unocc_idx]
        # non-redundant orbital rotations (q > p)
        inact, act, virt = set(self.inactive_idx), set(self.active_idx), set(self.virtual_idx)
        self._kappa, self._kappa_old = [], []
        kappa_idx, no_aa, no_aa_dag, redundant, hf_like = [], [], [], [], []
        for p in range(self.num_orbs):
            for q in range(p + 1, self.num_orbs):
                if (p in inact and q in inact) or (p in virt and q in virt):
                    redundant.append((p, q))
                    continue
                if not include_active_kappa and p in act and q in act:
                    redundant.append((p, q))
                    continue
                if not (p in act and q in act):
                    no_aa.append((p, q))
                    no_aa_dag.append((q, p))
                self._kappa.append(0.0)
                self._kappa_old.append(0.0)
                kappa_idx.append((p, q))
        occ_a, unocc_a = set(self.active_occ_idx), set(self.active_unocc_idx)
        for p in range(self.num_orbs):
            for q in range(p + 1, self.num_orbs):
                if (p in inact and (q in virt or q in unocc_a)) or (p in occ_a and q in virt):
                    hf_like.append((p, q))
        self.kappa_idx = np.array(kappa_idx, dtype=int)
        self.kappa_no_activeactive_idx = np.array(no_aa, dtype=int)
        self.kappa_no_activeactive_idx_dagger = np.array(no_aa_dag, dtype=int)
        self.kappa_redundant_idx = np.array(redundant, dtype=int)
        self.kappa_hf_like_idx = np.array(hf_like, dtype=int)
        self.ci_info = get_indexing(
            self.num_inactive_orbs, self.num_active_orbs, self.num_virtual_orbs,
            self.num_active_elec_alpha, self.num_active_elec_beta,
        )
        self.num_det = len(self.ci_info.idx2det)
        self._dev_space = None
        self._dev_ham = None
        self._dev_ham_key = None

    # ---- orbital rotations -------------------------------------------------
    @property
    def kappa(self) -> list[float]:
        return self._kappa.copy()

    @kappa.setter
    def kappa(self, k: list[float]) -> None:
        """Moves the expansion point (ups_wavefunction.py:321-334).  Setting
        the kappa already set, with no expansion point pending, changes
        nothing, so the integrals and everything built from them are kept
        (the optimisers set the same parameters for the energy and then the
        gradient)."""
        if list(k) == self._kappa and self._kappa_old == self._kappa:
            return
        self._h_mo = None
        self._g_mo = None
        self.__dict__["_mo_dev_cache"] = {}
        self._dev_ham = None  # built from the integrals just dropped
        self._energy_elec = None
        self._kappa = list(k).copy()
        self._c_mo = self.c_mo
        self._kappa_old = self.kappa

    @property
    def c_mo(self) -> np.ndarray:
        """C_old expm(-kappa) for the kappa moved since the last expansion point
        (ups_wavefunction.py:367-385), computed on the device."""
        if len(self.kappa) != 0:
            diff = np.array(self.kappa) - np.array(self._kappa_old)
            if np.max(np.abs(diff)) > 0.0:
                from . import orbitals

                kappa_mat = orbitals.kappa_matrix(diff, self.kappa_idx, self._c_mo.shape[1])
                return orbitals.rotate(self._c_mo, kappa_mat).cpu().numpy()
        return np.array(self._c_mo, dtype=np.float64, copy=True)

    def _ao_dev(self, which: str):
        """AO integrals on the device, uploaded once."""
        import torch

        cache = self.__dict__.setdefault("_ao_dev_cache", {})
        if which not in cache:
            src = self.int_gen.h_ao if which == "h" else self.int_gen.electron_electron_repulsion
            cache[which] = torch.as_tensor(np.ascontiguousarray(src, dtype=np.float64)).to("cuda")
        return cache[which]

    def _mo_dev(self, which: str):
        """MO integrals (C^T h C, or the 4-index g transform) on the device, cached until kappa moves."""
        from . import orbitals

        cache = self.__dict__.setdefault("_mo_dev_cache", {})
        if cache.get(which) is None:
            if which == "h":
                cache["h"] = orbitals.ao2mo(self.c_mo, self._ao_dev("h"), None)[0]
            else:
                cache["g"] = orbitals.ao2mo(self.c_mo, None, self._ao_dev("g"))[1]
        return cache[which]

    @property
    def h_mo(self) -> np.ndarray:
        if self._h_mo is None:
            self._h_mo = self._mo_dev("h").cpu().numpy()
        return self._h_mo

    @property
    def g_mo(self) -> np.ndarray:
        if self._g_mo is None:
            self._g_mo = self._mo_dev("g").cpu().numpy()
        return self._g_mo

    def _orbital_gradient(self, rdm1: np.ndarray, rdm2: np.ndarray) -> np.ndarray:
        """get_orbital_gradient (density_matrix.py:181-230) on the device integrals."""
        from . import orbitals

        return orbitals.orbital_gradient(self._mo_dev("h"), self._mo_dev("g"), self.kappa_idx, self.num_inactive_orbs,
                                         self.num_active_orbs, rdm1, rdm2)

    # ---- device pieces -----------------------------------------------------
    def _space(self):
        if self._dev_space is None:
            from . import engine

            self._dev_space = engine.space_for(self.num_active_orbs, self.num_active_elec_alpha, self.num_active_elec_beta)
        return self._dev_space

    def _csf_device(self):
        """csf_coeffs on the device (engine convention), converted once per
        array: the reference builds csf_coeffs in its constructor and never
        changes it; assigning a new array converts again."""
        c = self.__dict__.get("_csf_dev_cached")
        if c is None or c[0] is not self.csf_coeffs:
            c = (self.csf_coeffs, self._space().from_reference(self.csf_coeffs))
            self.__dict__["_csf_dev_cached"] = c
        return c[1]

    def _hamiltonian(self):
        """Folded active-space Hamiltonian on the device, rebuilt when h_mo/g_mo change."""
        from . import engine
        from .operators import folded_integrals

        h, g = self.h_mo, self.g_mo
        key = (id(h), id(g))
        if self._dev_ham is None or self._dev_ham_key != key:
            e_core, h_eff, g_act = folded_integrals(h, g, self.num_inactive_orbs, self.num_active_orbs)
            self._dev_ham = engine.Hamiltonian(self._space(), h_eff, g_act, e_core)
            self._dev_ham_key = key
        return self._dev_ham

    # ---- optimisation drivers (ups_wavefunction.py:799-1017, ucc_wavefunction.py:723-965) --
    def _print_parameter_info(self, orbital_optimization: bool) -> None:
        print("### Parameters information:")
        if orbital_optimization:
            print(f"### Number kappa: {len(self.kappa)}")
        print(f"### Number theta: {len(self._thetas)}")

    def _extra_options(self, optimizer_name: str) -> dict | None:
        """Optimizer options beyond fun/jac; RotoSolve needs them (the UPS class
        provides them, ups_wavefunction.py:853-861, 997-1005)."""
        return None

    def _reset_kappa_after_two_step(self) -> None:
        for i in range(len(self.kappa)):
            self._kappa[i] = 0.0
            self._kappa_old[i] = 0.0

    def run_wf_optimization_2step(self, optimizer_name: str, orbital_optimization: bool = False, tol: float = 1e-10,
                                  maxiter: int = 1000, is_silent_subiterations: bool = False) -> None:
        self._print_parameter_info(orbital_optimization)
        e_old = 1e12
        e_new = None
        print("Full optimization")
        print("Iteration # | Iteration time [s] | Electronic energy [Hartree] | Energy measurement #")
        for full_iter in range(int(maxiter)):
            t0 = time.time()
            if not is_silent_subiterations:
                print("--------Ansatz optimization")
                print("--------Iteration # | Iteration time [s] | Electronic energy [Hartree] | Energy measurement #")
            opt = Optimizers(
                partial(self._calc_energy_optimization, theta_optimization=True, kappa_optimization=False),
                optimizer_name,
                grad=partial(self._calc_gradient_optimization, theta_optimization=True, kappa_optimization=False),
                maxiter=maxiter, tol=tol, is_silent=is_silent_subiterations,
                energy_eval_callback=lambda: self.num_energy_evals,
            )
            self._old_opt_parameters = np.zeros_like(self.thetas) + 10**20
            self._E_opt_old = 0.0
            res = opt.minimize(self.thetas, extra_options=self._extra_options(optimizer_name))
            self.thetas = res.x.tolist()
            if orbital_optimization and len(self.kappa) != 0:
                if not is_silent_subiterations:
                    print("--------Orbital optimization")
                    print("--------Iteration # | Iteration time [s] | Electronic energy [Hartree] | Energy measurement #")
                opt = Optimizers(
                    partial(self._calc_energy_optimization, theta_optimization=False, kappa_optimization=True),
                    "l-bfgs-b",
                    grad=partial(self._calc_gradient_optimization, theta_optimization=False, kappa_optimization=True),
                    maxiter=maxiter, tol=tol, is_silent=is_silent_subiterations,
                    energy_eval_callback=lambda: self.num_energy_evals,
                )
                self._old_opt_parameters = np.zeros(len(self.kappa_idx)) + 10**20
                self._E_opt_old = 0.0
                res = opt.minimize([0.0] * len(self.kappa_idx))
                self._reset_kappa_after_two_step()
            else:
                e_new = res.fun
                if orbital_optimization and len(self.kappa) == 0:
                    print("WARNING: No orbital optimization performed, because there is no non-redundant orbital parameters.")
                break
            e_new = res.fun
            print(f"{str(full_iter + 1).center(11)} | {f'{time.time() - t0:7.2f}'.center(18)} | "
                  f"{f'{e_new:3.12f}'.center(27)} | {str(self.num_energy_evals).center(11)}")
            if abs(e_new - e_old) < tol:
                break
            e_old = e_new
        self._energy_elec = e_new

    def run_wf_optimization_1step(self, optimizer_name: str, orbital_optimization: bool = False, tol: float = 1e-10,
                                  maxiter: int = 1000) -> None:
        self._print_parameter_info(orbital_optimization)
        if optimizer_name.lower() == "rotosolve" and orbital_optimization and len(self.kappa) != 0:
            raise ValueError("Cannot use RotoSolve together with orbital optimization in the one-step solver.")
        print("--------Iteration # | Iteration time [s] | Electronic energy [Hartree] | Energy measurement #")
        theta_opt = not orbital_optimization or len(self.thetas) > 0
        kappa_opt = bool(orbital_optimization)
        energy = partial(self._calc_energy_optimization, theta_optimization=theta_opt, kappa_optimization=kappa_opt)
        gradient = partial(self._calc_gradient_optimization, theta_optimization=theta_opt, kappa_optimization=kappa_opt)
        if orbital_optimization:
            parameters = self.kappa + self.thetas if len(self.thetas) > 0 else self.kappa
        else:
            parameters = self.thetas
        opt = Optimizers(energy, optimizer_name, grad=gradient, maxiter=maxiter, tol=tol,
                         energy_eval_callback=lambda: self.num_energy_evals)
        self._old_opt_parameters = np.zeros_like(parameters) + 10**20
        self._E_opt_old = 0.0
        res = opt.minimize(parameters, extra_options=self._extra_options(optimizer_name))
        if orbital_optimization:
            self.thetas = res.x[len(self.kappa):].tolist()
            for i in range(len(self.kappa)):
                self._kappa[i] = 0.0
                self._kappa_old[i] = 0.0
        else:
            self.thetas = res.x.tolist()
        self._energy_elec = res.fun
