"""Build libfastgpr.so for gfx950 with hipcc (no cmake, no JIT cache).

The shared object lands next to the Python package
(``fast-gpr-fwi_amd/fastgpr/libfastgpr.so``) so it travels with the source
tree.  Run ``python fast-gpr-fwi_amd/csrc/build.py`` or call ``build()``.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
PKG = HERE.parent / "fastgpr"
REPO = HERE.parent.parent
INCLUDE = REPO / "include"
OUT = PKG / "libfastgpr.so"
SOURCES = [HERE / "fastgpr.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # Reference operation order: no fused multiply-add contraction.
    "-ffp-contract=off",
    # Packed FP32 instructions are kept out of the device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build(verbose: bool = False) -> Path:
    cmd = [HIPCC, *FLAGS, f"-I{INCLUDE}", "-o", str(OUT) + ".tmp", *map(str, SOURCES)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(str(OUT) + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
