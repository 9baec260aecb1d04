"""Builds libpqgpu.so (HIP for gfx950) in-tree with hipcc.

    python -m polyquant_amd.build        # or build() from Python

Every HIP source is compiled with packed-FP32 instructions disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "libpqgpu.so")
OBJDIR = os.path.join(HERE, "_build")

SOURCES = ["pqg_ctx.cpp", "pqg_ci.hip", "pqg_tensor.hip", "pqg_linkgemm.hip", "pqg_linkgemm_b3.hip", "pqg_davidson.hip", "pqg_ao2mo.hip", "pqg_rdm.hip",
           "pqg_spin.hip", "pqg_serve.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PQG_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++20", f"--offload-arch={ARCH}", "-I", os.path.join(ROOT, "include"),
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def _compile(src: str) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", path, "-o", obj]
    subprocess.run(cmd, check=True)
    return obj


def build(verbose: bool = True) -> str:
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", OUT + ".tmp"] + objs + ["-lpthread"]
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
