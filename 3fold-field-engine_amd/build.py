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
import tempfile
from concurrent.futures import ThreadPoolExecutor
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


def build(verbose=False, jobs=None, extra_flags=(), extra_sources=(), out=None, shared=True):
    """Compiles each source to an object in parallel (separate translation
    units, as one hipcc call would make them), then links the library.

    extra_flags / extra_sources / out / shared: a variant build (the host
    sanitizer run of the tests links a driver program against all sources)."""
    jobs = jobs or min(8, os.cpu_count() or 1)
    out = Path(out) if out else LIB
    tmp = out.with_name(out.name + ".tmp")
    with tempfile.TemporaryDirectory(prefix="fe_hip_build_") as objdir:
        objs, cmds = [], []
        for src in [*sources(), *map(str, extra_sources)]:
            obj = str(Path(objdir) / (Path(src).stem + ".o"))
            objs.append(obj)
            cmds.append([HIPCC, *[f for f in FLAGS if f != "-shared"], *extra_flags, "-c", src, "-o", obj])

        def run(cmd):
            if verbose:
                print(" ".join(cmd), file=sys.stderr)
            subprocess.run(cmd, check=True)

        with ThreadPoolExecutor(max_workers=jobs) as pool:
            for f in [pool.submit(run, c) for c in cmds]:
                f.result()
        run([HIPCC, f"--offload-arch={ARCH}", *(["-shared"] if shared else []), "-fPIC", *extra_flags, *objs,
             "-o", str(tmp)])
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
