"""Build libcrysplight_hip.so (gfx950) in-tree with hipcc.

Usage: python -m crysplight_amd.build [--verbose]

Flags that the transport's bit-parity depends on:
  -ffp-contract=off                        no fused multiply-add contraction (Julia
                                           evaluates a*b+c as two rounded ops)
  -fhip-fp32-correctly-rounded-divide-sqrt IEEE division and sqrt in binary32
  -fno-fast-math, denormals kept           (HIP default for gfx9)
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 instructions
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
LIB = HERE / "libcrysplight_hip.so"
ARCH = os.environ.get("CRYSP_OFFLOAD_ARCH", "gfx950")

SOURCES = ["transport.hip", "surrogate.hip", "gamma.cpp"]

COMMON = [
    "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def hipcc() -> str:
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(p).exists():
        raise RuntimeError("hipcc not found: the HIP extension cannot be built")
    return p


def build(verbose: bool = False) -> Path:
    """Compile every HIP/C++ source of the library and link the shared object."""
    objdir = HERE / "_build"
    objdir.mkdir(exist_ok=True)
    objs = []
    for src in SOURCES:
        path = CSRC / src
        if not path.exists():
            continue
        obj = objdir / (src + ".o")
        if path.suffix == ".hip":
            cmd = [hipcc(), *COMMON, *DEVICE, "-x", "hip", "-c", str(path), "-o", str(obj)]
        else:
            cmd = [hipcc(), *COMMON, "-c", str(path), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compiling {src} failed:\n{r.stdout}\n{r.stderr}")
        if verbose and r.stderr.strip():
            print(r.stderr, file=sys.stderr)
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs,
           "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"linking failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose: bool = False) -> None:
    """Build the CPU oracle (test infrastructure) under oracle/."""
    r = subprocess.run(["make", "-B", "-s", "-C", str(ROOT / "oracle")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    if verbose and r.stdout.strip():
        print(r.stdout)


if __name__ == "__main__":
    v = "--verbose" in sys.argv
    print(build(verbose=v))
