"""Module-level state functions with the reference's signatures
(operator_state_algebra.py:472-2976), executed by the HIP engine.

Inputs and outputs are host numpy vectors in the reference's index order and
sign convention; every function returns a new array and never mutates its
inputs, like the reference.  Internally each call converts to the engine's
sign convention on the device, runs the kernels and converts back.
"""

from __future__ import annotations

from collections.abc import Sequence

import numpy as np
import torch

from . import engine as E
from .ci_spaces import CI_Info
from .fermionic_operator import FermionicOperator
from .util import UccStructure, UpsStructure


def _space(ci: CI_Info) -> E.Space:
    if ci.space_extension_offset != 0:
        raise NotImplementedError("extended CI spaces (linear response) are not supported by the engine")
    return E.space_for(ci.num_active_orbs, ci.num_active_elec_alpha, ci.num_active_elec_beta)


_PROGRAMS: dict = {}


def _program(sp: E.Space, layout) -> E.Program:
    key = (id(sp), tuple(layout.excitation_operator_type), tuple(tuple(x) for x in layout.excitation_indices))
    prog = _PROGRAMS.get(key)
    if prog is None:
        prog = E.Program(sp, layout.excitation_operator_type, layout.excitation_indices)
        if len(_PROGRAMS) > 64:
            _PROGRAMS.clear()
        _PROGRAMS[key] = prog
    return prog


_OPERATORS: dict = {}


def _operator(sp: E.Space, op: FermionicOperator, do_unsafe: bool) -> E.Operator:
    """The compiled form of an (already folded) operator, reused across calls:
    its second application on the device goes through the operator's
    sparsity pattern instead of the term scan (linear response applies the
    same operators many times).  Keyed by the operator's strings and
    coefficients; at most 64 are kept."""
    key = (id(sp), bool(do_unsafe), tuple(op.operators.items()))
    O = _OPERATORS.get(key)
    if O is None:
        O = E.Operator(sp, op.operators.items(), drop_outside=do_unsafe)
        if len(_OPERATORS) >= 64:
            _OPERATORS.pop(next(iter(_OPERATORS)))
        _OPERATORS[key] = O
    return O


def _ucc_program(sp: E.Space, layout: UccStructure) -> E.Program:
    types = ["sa_single_ucc" if t == "sa_single" else t for t in layout.excitation_operator_type]

    class _L:
        excitation_operator_type = types
        excitation_indices = layout.excitation_indices

    return _program(sp, _L)


def _to_device(sp: E.Space, state) -> torch.Tensor:
    v = np.asarray(state, dtype=np.float64)
    if v.shape != (sp.dim,):
        raise ValueError(f"state has shape {v.shape}, expected ({sp.dim},)")
    return sp.from_reference(v)


def _to_host(sp: E.Space, t: torch.Tensor) -> np.ndarray:
    return sp.to_reference_numpy(t)


def construct_ups_state(state, ci_info: CI_Info, thetas, ups_struct: UpsStructure, dagger: bool = False) -> np.ndarray:
    sp = _space(ci_info)
    if len(thetas) != ups_struct.n_params:
        raise ValueError(f"Expected {ups_struct.n_params} thetas got {len(thetas)}")
    prog = _program(sp, ups_struct)
    out = prog.state_prep(np.asarray(thetas, dtype=np.float64), _to_device(sp, state), dagger=dagger)
    return _to_host(sp, out)


def propagate_unitary(state, idx: int, ci_info: CI_Info, thetas, ups_struct: UpsStructure) -> np.ndarray:
    sp = _space(ci_info)
    prog = _program(sp, ups_struct)
    psi = _to_device(sp, state)
    prog.apply_gate(idx, np.asarray(thetas, dtype=np.float64), psi)
    return _to_host(sp, psi)


def get_grad_action(state, idx: int, ci_info: CI_Info, ups_struct: UpsStructure) -> np.ndarray:
    sp = _space(ci_info)
    prog = _program(sp, ups_struct)
    return _to_host(sp, prog.grad_action(idx, _to_device(sp, state)))


def get_ucc_T(thetas, ucc_struct: UccStructure, offset: int = 0) -> FermionicOperator:
    """T(theta) as a FermionicOperator (operator_state_algebra.py:899-960)."""
    from .operators import G1, G2, G3, G4, G5, G6, G1_sa, G2_sa

    T = FermionicOperator({})
    for t, idx, th in zip(ucc_struct.excitation_operator_type, ucc_struct.excitation_indices, thetas):
        if abs(th) < 1e-28:
            continue
        th = float(th)
        if t == "sa_single":
            T += th * G1_sa(*(np.array(idx) + offset), True)
        elif t.startswith("sa_double_"):
            T += th * G2_sa(*(np.array(idx) + offset), int(t[-1]), True)
        else:
            fn = {"single": G1, "double": G2, "triple": G3, "quadruple": G4, "quintuple": G5, "sextuple": G6}.get(t)
            if fn is None:
                raise ValueError(f"Got unknown excitation type, {t}")
            T += th * fn(*(np.array(idx) + 2 * offset), True)
    return T


def construct_ucc_state(state, ci_info: CI_Info, thetas, ucc_struct: UccStructure, dagger: bool = False) -> np.ndarray:
    """exp(T(theta)) state, matrix-free on the GPU (operator_state_algebra.py:870-896)."""
    sp = _space(ci_info)
    prog = _ucc_program(sp, ucc_struct)
    th = np.asarray(thetas, dtype=np.float64)
    out = E.ucc_expm(prog, th, _to_device(sp, state), -1.0 if dagger else 1.0)
    return _to_host(sp, out)


def _fold(op: FermionicOperator, ci: CI_Info, do_folding: bool) -> FermionicOperator:
    if do_folding:
        return op.get_folded_operator(ci.num_inactive_orbs, ci.num_active_orbs, ci.num_virtual_orbs)
    return op


def _compile_list(operators, ci: CI_Info, sp: E.Space, thetas, wf_struct, do_folding, do_unsafe, ups_only=False):
    """Device steps psi -> psi for an operator list, rightmost first; the
    argument errors of the reference are raised before any device work.
    ups_only: "U"/"Ud" accept only a UpsStructure (the _SA functions,
    operator_state_algebra.py:684-700)."""
    steps = []
    for op in operators[::-1]:
        if isinstance(op, str):
            if op not in ("U", "Ud"):
                raise ValueError(f"Unknown str operator, expected ('U', 'Ud') got {op}")
            if ups_only and not isinstance(wf_struct, UpsStructure):
                raise TypeError(f"Got unknown wave function structure type, {type(wf_struct)}")
            if thetas is None:
                raise ValueError("theta must be different from None")
            th = np.asarray(thetas, dtype=np.float64)
            dag = op == "Ud"
            if isinstance(wf_struct, UpsStructure):
                prog = _program(sp, wf_struct)
                steps.append(lambda psi, prog=prog, th=th, dag=dag: prog.state_prep(th, psi, dagger=dag))
            elif isinstance(wf_struct, UccStructure):
                prog = _ucc_program(sp, wf_struct)
                steps.append(lambda psi, prog=prog, th=th, dag=dag: E.ucc_expm(prog, th, psi, -1.0 if dag else 1.0))
            else:
                raise TypeError(f"Got unknown wave function structure type, {type(wf_struct)}")
        else:
            folded = _fold(op, ci, do_folding)
            steps.append(_operator(sp, folded, do_unsafe).apply)
    return steps


def _apply_list(operators, psi: torch.Tensor, ci: CI_Info, sp: E.Space, thetas, wf_struct, do_folding, do_unsafe):
    for step in _compile_list(operators, ci, sp, thetas, wf_struct, do_folding, do_unsafe):
        psi = step(psi)
    return psi


def propagate_state(operators: list, state, ci_info: CI_Info, thetas: Sequence[float] | None = None,
                    wf_struct: UpsStructure | UccStructure | None = None, do_folding: bool = True,
                    do_unsafe: bool = False) -> np.ndarray:
    """Apply a list of FermionicOperators and/or "U"/"Ud" right to left."""
    if len(operators) == 0:
        return np.array(state, dtype=np.float64, copy=True)
    sp = _space(ci_info)
    psi = _apply_list(operators, _to_device(sp, state), ci_info, sp, thetas, wf_struct, do_folding, do_unsafe)
    return _to_host(sp, psi)


def expectation_value(bra, operators: list, ket, ci_info: CI_Info, thetas=None, wf_struct=None,
                      do_folding: bool = True, do_unsafe: bool = False) -> float:
    val = np.asarray(bra, dtype=np.float64) @ propagate_state(operators, ket, ci_info, thetas, wf_struct,
                                                               do_folding=do_folding, do_unsafe=do_unsafe)
    return float(val)


def build_operator_matrix(op: FermionicOperator, ci_info: CI_Info, do_unsafe: bool = False) -> np.ndarray:
    """Dense matrix of an active-space operator (operator_state_algebra.py:413-469),
    built in one launch on the device (sq_operator_matrix)."""
    sp = _space(ci_info)
    return _operator(sp, op, do_unsafe).matrix().cpu().numpy()


# ---- state-averaged (leading state axis) forms -------------------------------
# operator_state_algebra.py:282-410, 633-867, 1415-1864, 2312-2976: the same
# operations on every row of a (n_states, dim) array.  The array crosses to
# the device and back once per call; each row is propagated there with the
# single-state kernels (no host synchronisation between rows); programs and
# operators are built once per call.


def _rows(sp: E.Space, state) -> list[torch.Tensor]:
    """Rows of a (n_states, dim) host array on the device in the engine
    convention: one host-to-device copy for the whole array."""
    a = np.asarray(state, dtype=np.float64)
    if a.ndim != 2 or a.shape[1] != sp.dim:
        raise ValueError(f"states have shape {a.shape}, expected (n_states, {sp.dim})")
    X = torch.as_tensor(np.ascontiguousarray(a)).to("cuda")
    return [sp.from_reference(X[k]) for k in range(X.shape[0])]


def _rows_to_host(sp: E.Space, rows: list[torch.Tensor]) -> np.ndarray:
    """Device rows back to a host (n_states, dim) array in the reference
    convention: one device-to-host copy."""
    if not rows:
        return np.zeros((0, sp.dim))
    return torch.stack([sp.to_reference(r) for r in rows]).cpu().numpy()


def propagate_state_SA(operators: list, state, ci_info: CI_Info, thetas: Sequence[float] | None = None,
                       wf_struct: UpsStructure | None = None, do_folding: bool = True,
                       do_unsafe: bool = False) -> np.ndarray:
    """propagate_state on every row (operator_state_algebra.py:633-781);
    "U"/"Ud" take a UpsStructure only, as there."""
    if len(operators) == 0:
        return np.array(state, dtype=np.float64, copy=True)
    sp = _space(ci_info)
    rows = _rows(sp, state)
    steps = _compile_list(operators, ci_info, sp, thetas, wf_struct, do_folding, do_unsafe, ups_only=True)
    out = []
    for psi in rows:
        for step in steps:
            psi = step(psi)
        out.append(psi)
    return _rows_to_host(sp, out)


def expectation_value_SA(bra, operators: list, ket, ci_info: CI_Info, thetas: Sequence[float] | None = None,
                         wf_struct: UpsStructure | None = None, do_folding: bool = True) -> float:
    """(1/n_states) sum_k <bra_k|O|ket_k> (operator_state_algebra.py:827-867)."""
    op_ket = propagate_state_SA(operators, ket, ci_info, thetas, wf_struct, do_folding=do_folding)
    b = np.asarray(bra, dtype=np.float64)
    return float(np.einsum("kd,kd->", b, op_ket) / len(b))


def construct_ups_state_SA(state, ci_info: CI_Info, thetas: Sequence[float], ups_struct: UpsStructure,
                           dagger: bool = False) -> np.ndarray:
    """U(theta) (or U^+) on every row (operator_state_algebra.py:1415-1864)."""
    sp = _space(ci_info)
    if len(thetas) != ups_struct.n_params:
        raise ValueError(f"Expected {ups_struct.n_params} thetas got {len(thetas)}")
    rows = _rows(sp, state)
    prog = _program(sp, ups_struct)
    th = prog.theta_tensor(np.asarray(thetas, dtype=np.float64))
    return _rows_to_host(sp, [prog.state_prep(th, r, dagger=dagger) for r in rows])


def propagate_unitary_SA(state, idx: int, ci_info: CI_Info, thetas: Sequence[float],
                         ups_struct: UpsStructure) -> np.ndarray:
    """Gate idx on every row (operator_state_algebra.py:2312-2754)."""
    sp = _space(ci_info)
    rows = _rows(sp, state)
    prog = _program(sp, ups_struct)
    th = prog.theta_tensor(np.asarray(thetas, dtype=np.float64))
    for r in rows:
        prog.apply_gate(idx, th, r)
    return _rows_to_host(sp, rows)


def get_grad_action_SA(state, idx: int, ci_info: CI_Info, ups_struct: UpsStructure) -> np.ndarray:
    """T_idx on every row (operator_state_algebra.py:2868-2976)."""
    sp = _space(ci_info)
    rows = _rows(sp, state)
    prog = _program(sp, ups_struct)
    return _rows_to_host(sp, [prog.grad_action(idx, r) for r in rows])
