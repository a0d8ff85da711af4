"""Build recipe for libq3x.so (HIP for gfx950, host C++17), in-tree.

Usage: python -m q3_amd.build   (or q3_amd.build.build() from __graft_entry__.build()).
Objects go to q3_amd/_build/, the library to q3_amd/libq3x.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libq3x.so")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result",
    # packed fp32 vector ops are kept out of every kernel on purpose
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# per-file extras: the setup code must not contract a*b+c into FMA (bit-exact G maps)
EXTRA = {
    "setup.cpp": ["-ffp-contract=off"],
    # subspace eigensolver: reassociation lets the reductions vectorise (accuracy is to rounding)
    "dense.cpp": ["-fopenmp", "-mavx2", "-mfma", "-fassociative-math", "-fno-signed-zeros", "-fno-trapping-math",
                  "-fno-math-errno"],
    "davidson.cpp": ["-fopenmp"],
}
SOURCES = ["setup.cpp", "kb_projectors.cpp", "api.cpp", "hlocal.hip", "plane_tl.hip", "zgemm.hip", "aux_kernels.hip",
           "subspace_eig.hip", "potential.hip", "noalloc.hip", "gamma_kernels.hip", "kb_device.hip", "dense.cpp", "davidson.cpp", "batch.cpp", "server.cpp",
           "memory_estimate.cpp"]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(OUT, src + ".o")
    cmd = [_hipcc(), *HIP_FLAGS, *EXTRA.get(src, []), "-x", "hip", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OUT, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "--offload-arch=gfx950", "-shared", "-fopenmp", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
