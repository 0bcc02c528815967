"""ctypes binding of the C ABI in include/sqhip.h (libsqhip.so, built in-tree).

torch is imported first so that the library's HIP runtime dependency resolves
to the runtime torch already loaded: one HIP runtime per process.  There is
no fallback: if the library is missing or fails to load, every engine call
raises.
"""

from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (must precede the library load, see module docstring)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsqhip.so")

SQ_OK = 0
SQ_ERR_ARG = -1
SQ_ERR_HIP = -2
SQ_ERR_SPACE = -3
SQ_ERR_TYPE = -4

GATE_KIND = {
    "sa_single": 1,
    "single": 2,
    "double": 3,
    "triple": 4,
    "quadruple": 5,
    "quintuple": 6,
    "sextuple": 7,
    "sa_double_1": 11,
    "sa_double_2": 12,
    "sa_double_3": 13,
    "sa_double_4": 14,
    "sa_double_5": 15,
    "sa_single_ucc": 21,
}


class GateDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("n_idx", ctypes.c_int32), ("idx", ctypes.c_int32 * 12)]


class OpString(ctypes.Structure):
    _fields_ = [
        ("coef", ctypes.c_double),
        ("n_ops", ctypes.c_int32),
        ("spin_orb", ctypes.c_int32 * 24),
        ("is_creator", ctypes.c_int32 * 24),
    ]


_lib = None

vp = ctypes.c_void_p
i32 = ctypes.c_int
i64 = ctypes.c_int64
dp = ctypes.c_void_p  # device pointers are passed as integers

_SIGNATURES = {
    "sq_last_error": ([], ctypes.c_char_p),
    "sq_version": ([], ctypes.c_char_p),
    "sq_space_create": ([i32, i32, i32, ctypes.POINTER(vp)], i32),
    "sq_space_destroy": ([vp], i32),
    "sq_space_dim": ([vp], i64),
    "sq_space_nstrings": ([vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)], i32),
    "sq_space_copy_strings": ([vp, vp, vp], i32),
    "sq_space_idx2det": ([vp, vp, vp], i32),
    "sq_state_from_reference": ([vp, dp, dp, vp], i32),
    "sq_state_to_reference": ([vp, dp, dp, vp], i32),
    "sq_program_create": ([vp, ctypes.POINTER(GateDesc), i32, ctypes.POINTER(vp)], i32),
    "sq_program_destroy": ([vp], i32),
    "sq_program_ngates": ([vp], i32),
    "sq_program_set_tile_shape": ([vp, i32, i32], i32),
    "sq_state_prep": ([vp, vp, dp, dp, dp, i32, dp, vp], i32),
    "sq_apply_gate": ([vp, vp, i32, dp, dp, dp, vp], i32),
    "sq_grad_action": ([vp, vp, i32, dp, dp, vp], i32),
    "sq_hamiltonian_create": ([vp, vp, vp, ctypes.c_double, ctypes.POINTER(vp)], i32),
    "sq_hamiltonian_destroy": ([vp], i32),
    "sq_work_bytes": ([vp, i32], i64),
    "sq_sigma": ([vp, vp, dp, dp, vp, dp, vp], i32),
    "sq_sigma_noalloc": ([vp, vp, dp, dp, dp, vp], i32),
    "sq_sigma_configure": ([vp, i32, i32, i32], i32),
    "sq_energy_grad_adjoint": ([vp, vp, vp, dp, dp, vp, vp, dp, vp], i32),
    "sq_energy_grad_plain": ([vp, vp, vp, dp, dp, vp, vp, dp, vp], i32),
    "sq_energy_grad_reference_sweep": ([vp, vp, vp, dp, dp, vp, vp, dp, vp], i32),
    "sq_adjoint_reverse": ([vp, vp, dp, dp, dp, vp, dp, vp], i32),
    "sq_rdm12": ([vp, dp, vp, vp, dp, vp], i32),
    "sq_debug_gemm": ([dp, dp, dp, i32, i32, i32, i64, vp], i32),
    "sq_operator_create": ([vp, ctypes.POINTER(OpString), i32, i32, ctypes.POINTER(vp)], i32),
    "sq_operator_destroy": ([vp], i32),
    "sq_operator_apply": ([vp, vp, dp, dp, vp], i32),
    "sq_operator_matrix": ([vp, vp, i32, dp, vp], i32),
    "sq_ucc_work_bytes": ([vp, vp], i64),
    "sq_ucc_expm": ([vp, vp, dp, ctypes.c_double, dp, dp, dp, vp], i32),
    "sq_ucc_energy_grad": ([vp, vp, vp, dp, dp, vp, vp, dp, vp], i32),
    "sq_batch_create": ([i32, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp), i32, ctypes.POINTER(vp)], i32),
    "sq_batch_destroy": ([vp], i32),
    "sq_batch_num_params": ([vp], i64),
    "sq_batch_set_theta": ([vp, vp, vp], i32),
    "sq_batch_set_start_states": ([vp, vp, vp], i32),
    "sq_batch_configure": ([vp, i32, i32], i32),
    "sq_batch_run": ([vp, vp], i32),
    "sq_batch_run_graph": ([vp, vp], i32),
    "sq_batch_results": ([vp, vp, vp, vp], i32),
    "sq_profile_enable": ([i32], i32),
    "sq_profile_select": ([ctypes.c_uint32], i32),
    "sq_profile_nclasses": ([], i32),
    "sq_profile_class_name": ([i32], ctypes.c_char_p),
    "sq_profile_collect": ([i32, vp, vp, vp, vp], i32),
    "sq_orbital_work_bytes": ([i32, i32, i32, i32, i32], i64),
    "sq_orbital_rotation": ([i32, i32, dp, vp, dp, dp, vp], i32),
    "sq_ao2mo": ([i32, i32, dp, dp, dp, dp, dp, dp, vp], i32),
    "sq_orbital_gradient": ([i32, i32, i32, dp, dp, dp, dp, i32, vp, vp, dp, vp], i32),
}

# Entry points added by later translation units register themselves here.
EXTRA_SIGNATURES: dict = {}


def exported_symbols() -> list[str]:
    return sorted(list(_SIGNATURES) + list(EXTRA_SIGNATURES))


def lib():
    """The loaded engine library; raises if it is missing (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"HIP engine library not found at {LIB_PATH}; build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc, gfx950)"
            )
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in {**_SIGNATURES, **EXTRA_SIGNATURES}.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


class EngineError(RuntimeError):
    pass


def check(rc: int) -> None:
    if rc == SQ_OK:
        return
    msg = lib().sq_last_error().decode(errors="replace")
    if rc == SQ_ERR_SPACE:
        raise KeyError(msg)
    if rc in (SQ_ERR_ARG, SQ_ERR_TYPE):
        raise ValueError(msg)
    raise EngineError(f"sqhip error {rc}: {msg}")


def gate_descs(types, indices) -> ctypes.Array:
    arr = (GateDesc * max(len(types), 1))()
    for k, (t, idx) in enumerate(zip(types, indices)):
        if t not in GATE_KIND:
            raise ValueError(f"Got unknown excitation type, {t}")
        arr[k].kind = GATE_KIND[t]
        arr[k].n_idx = len(idx)
        for j, x in enumerate(idx):
            arr[k].idx[j] = int(x)
    return arr
