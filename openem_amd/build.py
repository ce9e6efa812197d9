"""In-tree build of ``libopenem_hip.so`` for gfx950 (no JIT cache, so the .so travels with the
tree).  ``python -m openem_amd.build``."""

from __future__ import annotations

import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
SOURCES = [PKG / "csrc" / f for f in ("openem_hip.hip", "oem_pool.hip", "oem_reduce.hip",
                                                 "oem_adjoint.hip", "oem_host.cpp")]
OUT = PKG / "libopenem_hip.so"
ARCH = "gfx950"
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
         # no packed-FP32 instructions: hipcc may form v_pk_* from scalar float code
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         # no multiply-add contraction: every kernel rounds an expression the way it is written,
         # so the fused step, the two sweeps and the pool kernels agree bit for bit
         "-ffp-contract=off"]


def hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and Path(c).exists():
            return c
    raise RuntimeError("hipcc not found")


def build_library(verbose: bool = False) -> Path:
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), *FLAGS, "-o", str(tmp), *map(str, SOURCES)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{r.stderr[-4000:]}")
    if verbose and r.stderr.strip():
        print(r.stderr, file=sys.stderr)
    tmp.replace(OUT)
    return OUT


if __name__ == "__main__":
    print(build_library(verbose=True))
