"""Build the native library ``scx_amd/csrc/libscxpca.so`` for gfx950.

Run as ``python -m scx_amd.build``; ``__graft_entry__.build()`` calls
:func:`build`. hipcc cross-compiles without a GPU.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(CSRC, "libscxpca.so")
ARCH = os.environ.get("SCXPCA_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions stay off (v_pk_fma_f32 and friends)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build scx_amd)")


def build(verbose: bool = False) -> str:
    hipcc = _hipcc()
    obj_dir = os.path.join(CSRC, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    omega_o = os.path.join(obj_dir, "omega.o")
    scx_o = os.path.join(obj_dir, "scxpca.o")
    # Host-only RNG: plain g++, IEEE semantics, no FP contraction (the stream
    # must reproduce the reference's f64 arithmetic bit for bit).
    cmds = [
        ["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-c",
         os.path.join(CSRC, "omega.cpp"), "-o", omega_o],
        [hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, "scxpca.hip"), "-o", scx_o],
        [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB + ".tmp", scx_o, omega_o,
         "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"],
    ]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd, cwd=CSRC)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
