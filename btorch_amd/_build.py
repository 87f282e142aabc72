"""Build recipe for libbtx.so, the gfx950 HIP kernels of this package.

``build_library()`` runs hipcc directly (no JIT cache): the shared object is
written next to this file so it travels with the source tree.
"""

from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
CSRC = PKG_DIR / "csrc"
LIB_PATH = PKG_DIR / "libbtx.so"
SOURCES = ["btx_spmv.hip", "btx_step.hip", "btx_adjoint.hip", "btx_io.hip", "btx_build.hip"]

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=gfx950",
    "-fPIC",
    "-shared",
    # Bit-exact parity: every a*b+c rounds twice, as in the eager reference.
    "-ffp-contract=off",
    # Native global float atomics for the order-free accumulation mode.
    "-munsafe-fp-atomics",
    # No packed-fp32 instructions in generated code.
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def hipcc_path() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; install ROCm or set HIPCC")


def build_library(verbose: bool = False) -> Path:
    """Compile every HIP source into libbtx.so (always recompiles)."""
    srcs = [str(CSRC / s) for s in SOURCES if (CSRC / s).exists()]
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [hipcc_path(), *HIPCC_FLAGS, *srcs, "-o", str(tmp)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed ({proc.returncode}):\n{proc.stderr[-4000:]}")
    if verbose and proc.stderr:
        print(proc.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build_library(verbose=True))
