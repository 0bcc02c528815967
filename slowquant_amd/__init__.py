"""MI355X-native state-vector engine for SlowQuant's UPS and UCC wave functions.

Drop-in classes: WaveFunctionUPS, WaveFunctionSAUPS, WaveFunctionUCC (same constructors and
entry points as the reference); module functions with the reference's
signatures live in slowquant_amd.operator_state_algebra.  All state-vector
arithmetic runs in the in-tree HIP library libsqhip.so (gfx950).
"""

from .ci_spaces import CI_Info, get_indexing
from .fermionic_operator import FermionicOperator
from .util import UccStructure, UpsStructure


def __getattr__(name):
    # GPU-backed classes load torch and the HIP library only when first used.
    if name == "WaveFunctionUPS":
        from .ups_wavefunction import WaveFunctionUPS

        return WaveFunctionUPS
    if name == "WaveFunctionSAUPS":
        from .sa_ups_wavefunction import WaveFunctionSAUPS

        return WaveFunctionSAUPS
    if name == "WaveFunctionUCC":
        from .ucc_wavefunction import WaveFunctionUCC

        return WaveFunctionUCC
    raise AttributeError(name)


__all__ = ["CI_Info", "FermionicOperator", "UccStructure", "UpsStructure", "WaveFunctionSAUPS", "WaveFunctionUCC", "WaveFunctionUPS",
           "get_indexing"]
