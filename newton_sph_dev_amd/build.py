"""Build libnsph.so (gfx950) in-tree with hipcc.

    python -m newton_sph_dev_amd.build        # or: python newton_sph_dev_amd/build.py

The library is written next to this file so that it travels with the source tree.
Floating-point flags: no fast-math, FMA contraction off (the neighbour predicate and the
integrator must round exactly as written), IEEE divide/sqrt (hipcc default), and packed
FP32 instructions disabled.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
CSRC = PKG_DIR / "csrc"
LIB_PATH = PKG_DIR / "libnsph.so"
SOURCES = ["grid.hip", "sort.hip", "sph.hip", "mlp.hip", "contacts.hip", "shapes.hip"]
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libnsph.so)")


def flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-ffp-contract=off",
        "-fno-fast-math",
        "-Xclang",
        "-target-feature",
        "-Xclang",
        "-packed-fp32-ops",
        "-Wall",
        "-Wno-unused-function",
        "-Wno-unused-variable",
        "-Wno-unused-result",
    ]


def build(verbose: bool = False) -> Path:
    srcs = [str(CSRC / s) for s in SOURCES]
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [hipcc(), *flags(), "-o", str(tmp), *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed with exit code {proc.returncode}")
    if verbose and proc.stderr.strip():
        # the host pass warns about the packed-fp32 target feature; keep real warnings visible
        lines = [ln for ln in proc.stderr.splitlines() if "packed-fp32-ops" not in ln]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
