"""Compile the HIP library libmpm_hip.so in-tree for gfx950 with hipcc.

The library is built next to this file so that it travels with the source
tree; nothing is JIT-compiled or cached outside the repository.
"""
from __future__ import annotations

import os
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
INCLUDE = os.path.join(REPO_DIR, "include")
LIB_NAME = "libmpm_hip.so"
LIB_PATH = os.path.join(PKG_DIR, LIB_NAME)

SOURCES = ["mpm_kernels.hip", "mpm_noalloc.hip", "mpm_reduced.hip", "mpm_binned.hip", "mpm_ordered.hip", "mpm_adjoint.hip", "crom_mlp.hip", "crom_encoder.hip"]
HEADERS = ["mpm_common.h", "mpm_device.h", "mpm_stencil.h", "mpm_ordered.h"]

# -packed-fp32-ops: keep hipcc from emitting v_pk_{add,mul,fma}_f32.
# -munsafe-fp-atomics: hardware global_atomic_add_f32 instead of CAS loops.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def build(verbose: bool = False) -> str:
    """Compile every HIP source into LIB_PATH (always recompiles)."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", INCLUDE, "-I", CSRC, *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
