"""Builds libfrm.so (HIP, gfx950) in-tree.

    python -m theses_code_amd.build        # or __graft_entry__.build()

Every translation unit under csrc/ is compiled with hipcc for gfx950 only and
linked into theses_code_amd/libfrm.so next to this file.  Packed-FP32 ops are
disabled for device code (-target-feature -packed-fp32-ops).
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfrm.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(REPO, "include"), "-I" + CSRC,
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_cpp_example() -> list:
    """Host-only C++ programs that use include/frm_solvers.hpp (tests/cpp)."""
    outs = []
    for name in ("tg_dropin", "pso_batch", "timesum_dropin"):
        src = os.path.join(REPO, "tests", "cpp", name + ".cpp")
        out = os.path.join(REPO, "tests", "cpp", name)
        cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(REPO, "include"), src, "-o", out,
               "-L" + HERE, "-lfrm", "-Wl,-rpath," + HERE,
               "-Wl,-rpath,$ORIGIN/../../theses_code_amd"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"C++ example {name} build failed:\n{r.stderr}")
        outs.append(out)
    return outs


if __name__ == "__main__":
    build(verbose=True)
    build_cpp_example()
    sys.exit(0)
