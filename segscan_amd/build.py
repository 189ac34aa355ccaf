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
    "permute.hip",
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


# Host-sanitized build: the translation units that hold the C ABI's host logic (argument
# validation, planners, tuning tables, Matrix Market reader) are recompiled with
# AddressSanitizer and UndefinedBehaviorSanitizer on their host pass only (device code is
# unchanged); the kernel-only units are reused from _build/. Load it with the ASan runtime
# preloaded: see sanitized_env().
SAN_BUILD = PKG / "_build_san"
SAN_LIB = SAN_BUILD / "libsegscan_hip.so"
SAN_SOURCES = ["segscan_abi.cpp", "mtx_loader.cpp", "tune_table.hip", "noalloc.hip", "batch_kernels.hip", "permute.hip"]
SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=undefined", "-g"]


def _compile_san(src: str) -> Path:
    obj = SAN_BUILD / (Path(src).stem + ".o")
    cmd = [hipcc(), *CXXFLAGS, *SAN_FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build_sanitized(verbose: bool = True) -> Path:
    """libsegscan_hip.so with host-side ASan + UBSan in _build_san/ (the normal build first
    provides the kernel-only objects)."""
    if not all((BUILD / (Path(s).stem + ".o")).exists() for s in SOURCES):
        build(verbose)
    SAN_BUILD.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SAN_SOURCES)) as ex:
        san_objs = list(ex.map(_compile_san, SAN_SOURCES))
    objs = [o for o in (BUILD / (Path(s).stem + ".o") for s in SOURCES) if o.stem not in {Path(x).stem for x in SAN_SOURCES}]
    # host-side runtimes only: -fno-gpu-sanitize keeps the device link uninstrumented
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-fsanitize=address,undefined", "-fno-gpu-sanitize",
           "-shared-libsan", *map(str, objs), *map(str, san_objs), "-o", str(SAN_LIB)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(f"built {SAN_LIB}", file=sys.stderr)
    return SAN_LIB


def asan_runtime() -> Path:
    """The clang ASan runtime that a sanitized library needs preloaded into Python."""
    clang = Path(hipcc()).resolve().parent.parent / "llvm" / "bin" / "clang"
    if not clang.exists():
        clang = Path("/opt/rocm/llvm/bin/clang")
    out = subprocess.run([str(clang), "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True,
                         text=True).stdout.strip()
    return Path(out)


def sanitized_env(base: dict | None = None) -> dict:
    """Environment for a child Python that loads the sanitized library (leak checks off:
    the interpreter itself keeps its allocations until exit)."""
    env = dict(os.environ if base is None else base)
    env["LD_PRELOAD"] = str(asan_runtime())
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["SEGSCAN_LIB"] = str(SAN_LIB)
    return env


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
    if "--sanitize" in sys.argv[1:]:
        build_sanitized()
    else:
        build()
