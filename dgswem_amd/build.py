"""Build steps: the HIP library (gfx950), the CPU oracle and the reference binary.

    python -m dgswem_amd.build
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdgswem_hip.so")

# -ffp-contract=off: every product and sum is rounded on its own, as in the reference's x86-64
# build, so the GPU step reproduces dgswem_serial's arithmetic.  Packed FP32 instructions are
# switched off for all device code.
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
               "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wno-unused-command-line-argument"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build_hip() -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    _run([hipcc, *HIPCC_FLAGS, "-o", tmp, os.path.join(CSRC, "dg_capi.hip"),
          os.path.join(CSRC, "dg_setup.cpp")])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> str:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-B"])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_reference() -> bool:
    """oracle/_ref/dgswem_serial from /root/reference, and oracle/_ref/dgswem_gpu (the reference's
    Fortran host linked with the GPU step), only where the reference sources are present."""
    if not os.path.isdir(os.environ.get("DGSWEM_REFERENCE", "/root/reference")):
        return False
    _run(["bash", os.path.join(ROOT, "oracle", "build_ref.sh")])
    _run(["bash", os.path.join(ROOT, "oracle", "build_dropin.sh")])
    return True


def build_all() -> None:
    build_hip()
    build_oracle()
    build_reference()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
