"""Build libdiffstir_hip.so in-tree with hipcc for gfx950.

    python -m diffstir_amd.build          # or diffstir_amd.build.build()

The library is one translation unit (csrc/diffstir.hip includes the kernel files).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "diffstir.hip")
OUT = os.path.join(HERE, "libdiffstir_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-munsafe-fp-atomics",            # hardware float/double atomic add, no CAS loop
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",   # no v_pk_*_f32
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    f"-Wl,-rpath,{ROCM}/lib",
]


def sources():
    csrc = os.path.join(HERE, "csrc")
    out = [os.path.join(csrc, f) for f in sorted(os.listdir(csrc)) if f.endswith((".hip", ".h"))]
    out.append(os.path.join(os.path.dirname(HERE), "include", "diffstir.h"))
    return out


def build(verbose: bool = False) -> str:
    """Compile the library (always; the caller decides when) and return its path."""
    tmp = OUT + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for libdiffstir_hip.so")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
