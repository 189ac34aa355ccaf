"""Builds the HIP library (libsegscan_hip.so) for gfx950 in-tree.

Every translation unit under csrc/ is compiled with hipcc in parallel into
_build/, then linked into segscan_amd/libsegscan_hip.so next to this file.
Run as ``python -m segscan_amd.build`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libsegscan_hip.so"
INCLUDE = PKG.parent / "include"

SOURCES = [
    "pr_kernels.hip",
    "sr_f32.hip",
    "sr_bf16.hip",
    "sr_f16.hip",
    "bwd_kernels.hip",
    "rbsr_kernels.hip",
    "misc_kernels.hip",
    "segscan_abi.cpp",
    "noalloc.hip",
    "batch_kernels.hip",
    "tune_table.hip",
    "mtx_loader.cpp",
]

ARCH = os.environ.get("SEGSCAN_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: keep the compiler from forming v_pk_{add,mul,fma}_f32.
CXXFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build segscan_amd)")


def _compile(src: str) -> Path:
    obj = BUILD / (Path(src).stem + ".o")
    cmd = [hipcc(), *CXXFLAGS, "-c", str(CSRC / src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = True, jobs: int | None = None) -> Path:
    BUILD.mkdir(exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, os.cpu_count() or 1))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    build_vendor(verbose)
    return LIB


VENDOR_LIB = PKG / "libsegscan_vendor.so"


def build_vendor(verbose: bool = True) -> Path:
    """The hipSPARSE comparator (timing only), in a library of its own."""
    src = CSRC / "vendor_hipsparse.cpp"
    tmp = VENDOR_LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *CXXFLAGS, "-shared", str(src), "-lhipsparse", "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, VENDOR_LIB)
    if verbose:
        print(f"built {VENDOR_LIB}", file=sys.stderr)
    return VENDOR_LIB


if __name__ == "__main__":
    build()
