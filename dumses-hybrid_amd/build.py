"""Builds libdmx.so (gfx950) in-tree: dumses-hybrid_amd/build/libdmx.so.

    python dumses-hybrid_amd/build.py

The float64 translation unit is compiled with -ffp-contract=off, so its
kernels reproduce the reference's float64 arithmetic bit for bit; packed
FP32 instructions are disabled for all device code.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-result", "-I", os.path.join(PKG, "..", "include")]
UNITS = [
    ("dmx_f64.hip", ["-ffp-contract=off"]),
    ("dmx_f32.hip", ["-ffp-contract=fast", "-ffast-math", "-fno-finite-math-only"]),
    ("dmx_abi.cpp", []),
    ("dmx_mlp.hip", ["-ffp-contract=fast"]),
]


def _compile(unit, out=OUT, defines=()):
    src, extra = unit
    obj = os.path.join(out, src.rsplit(".", 1)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    cmd = [HIPCC, *COMMON, *extra, *defines, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False, out: str = OUT, defines=()) -> str:
    """Compiles and links libdmx.so into `out`; `defines` are extra -D
    options for the kernel sources (tuning variants)."""
    os.makedirs(out, exist_ok=True)
    with ThreadPoolExecutor(len(UNITS)) as ex:
        objs = list(ex.map(lambda u: _compile(u, out, tuple(defines)), UNITS))
    lib = os.path.join(out, "libdmx.so")
    tmp = lib + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    if verbose:
        print(lib)
    return lib


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
