"""Build the HIP engine library ``libsqhip.so`` in-tree with hipcc (gfx950)."""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsqhip.so")
SOURCES = [
    "sq_space.hip",
    "sq_gates.hip",
    "sq_sigma.hip",
    "sq_rowsigma.hip",
    "sq_grad.hip",
    "sq_profile.hip",
    "sq_tiles.hip",
    "sq_ucc.hip",
    "sq_batch.hip",
    "sq_orbital.hip",
    "sq_noalloc.hip",
]
ARCH = os.environ.get("SQ_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; the HIP engine cannot be built")


def build_library(verbose: bool = False) -> str:
    """Compile every engine source into libsqhip.so (always recompiles)."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [
        hipcc(),
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-Xclang",
        "-target-feature",
        "-Xclang",
        "-packed-fp32-ops",
        "-Wno-unused-result",
        "-I",
        os.path.join(HERE, "..", "include"),
        "-o",
        tmp,
        *srcs,
        "-L" + os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(hipcc()))), "lib"),
        "-lrocblas",  # the plain dense GEMM of the alpha-string Hamiltonian (sq_rowsigma.hip)
    ]
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stderr[-8000:]}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
