"""Builds libfe_hip.so (all HIP kernels and the C ABI) for gfx950 with hipcc.

    python 3fold-field-engine_amd/build.py

Floating point: no contraction and no fast-math, so every fp32/f64 operation
rounds once, as in the reference's native builds; fp32 division and sqrt stay
correctly rounded (hipcc's default for HIP).  Packed fp32 instructions are
disabled by target feature.
"""
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libfe_hip.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(str(p) for p in CSRC.glob("*.hip")) + sorted(str(p) for p in CSRC.glob("*.cpp"))


def build(verbose=False):
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, *FLAGS, *sources(), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
