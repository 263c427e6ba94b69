"""Build the HIP library in-tree: mgpoisson2d_amd/libmgp2d.so for gfx950.

Usage: python -m mgpoisson2d_amd.build   (or build() from __graft_entry__).
All HIP code is compiled with -ffp-contract=off (the kernels restate the
reference's FP64 arithmetic operand for operand) and without packed FP32 ops.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libmgp2d.so")
SOURCES = ["mgp_kernels.hip", "mgp_xline.hip", "mgp_flux.hip", "mgp_single.hip", "mgp_geometry.hip", "mgp_host.hip", "mgp_krylov.hip", "mgp_batch.hip", "mgp_rbtile.hip", "mgp_capi.hip", "mgp_noalloc.hip", "mgp_refine.hip", "mgp_fourth.hip", "mgp_gmres.hip"]
ARCH = os.environ.get("MGP_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wl,--no-undefined",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    cmd = [hipcc, *FLAGS, "-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (res.stdout + res.stderr).splitlines()
                    if "packed-fp32-ops" not in l and "feature flag" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed:\n{out}")
    if verbose and out.strip():
        print(out)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
