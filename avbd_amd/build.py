"""Build step of the avbd_amd package: compiles the HIP/C++ sources under
``avbd_amd/csrc`` for gfx950 into ``avbd_amd/_build/libavbd_amd.so`` (in-tree,
so the shared library travels with the repository to the GPU machine)."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT_DIR = PKG / "_build"
LIB = OUT_DIR / "libavbd_amd.so"
INCLUDE = PKG.parent / "include"

SOURCES = ["api.cpp", "host_graph.cpp", "contacts_host.cpp", "mlp_host.cpp", "vbd_kernels.hip", "lbvh_kernels.hip",
           "contact_kernels.hip", "mlp_kernels.hip", "lbvh_static.hip", "vbd_fp64.hip",
           "graph_kernels.hip", "recolour_kernels.hip"]
# FP64 Morton codes must round exactly like the reference's CPU code: no FMA
# contraction in the LBVH kernels.  The solver kernels contract only within
# one expression (-ffp-contract=on), so a device function rounds the same way
# wherever it is inlined: the fused and separate Chebyshev phases then
# produce identical bits.
EXTRA = {"lbvh_kernels.hip": ["-ffp-contract=off"], "lbvh_static.hip": ["-ffp-contract=off"], "contact_kernels.hip": ["-ffp-contract=off"],
         "vbd_kernels.hip": ["-ffp-contract=on"], "vbd_fp64.hip": ["-ffp-contract=on"]}

# -packed-fp32-ops: keep the compiler from emitting packed FP32 VALU ops
# (v_pk_{add,mul,fma}_f32) in any kernel of this library.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=fast",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build avbd_amd)")


def build(verbose: bool = False) -> Path:
    """Compile every source to an object, then link the shared library."""
    OUT_DIR.mkdir(exist_ok=True)
    cc = hipcc()
    objs = []
    for src in SOURCES:
        obj = OUT_DIR / (Path(src).stem + ".o")
        cmd = [cc, *HIP_FLAGS, *EXTRA.get(src, []), f"-I{INCLUDE}", "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
