"""Build libsolvix_hip.so in-tree with hipcc for gfx950.

Every source is compiled with hipcc --offload-arch=gfx950. The element and
recovery kernels are compiled with -ffp-contract=off so their arithmetic
keeps the reference's unfused order. All device code is built without packed
FP32 instructions (-target-feature -packed-fp32-ops).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsolvix_hip.so")
ARCH = os.environ.get("SVX_OFFLOAD_ARCH", "gfx950")

SOURCES = [
    ("assembly.hip", ["-ffp-contract=off"]),
    ("recover.hip", ["-ffp-contract=off"]),
    ("krylov.hip", []),
    ("block.hip", []),
    ("engine.cpp", []),
    ("pattern.cpp", []),
    ("model.cpp", []),
    ("api.cpp", []),
    ("batch_api.cpp", []),
    ("block_api.cpp", []),
    ("server.cpp", []),
    ("noalloc.cpp", []),
    ("deck.cpp", []),
]

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-I", INCLUDE,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src: str, extra) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    lang = ["-x", "hip"]
    cmd = [hipcc()] + lang + COMMON + extra + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [(s, e) for s, e in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda se: _compile(*se), srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
