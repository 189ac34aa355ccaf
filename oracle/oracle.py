"""numpy bindings of the CPU oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg, as the checker; the product package never imports it. See oracle.cpp for
the reference lines each function restates and how each is pinned.
"""
from __future__ import annotations

import ctypes
import subprocess
from ctypes import c_double, c_int32, c_int64, c_uint32, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-B", "-s", "-C", str(HERE)], check=True)
    return LIB


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        build()
    lib = ctypes.CDLL(str(LIB))
    P = c_void_p
    sig = {
        "oracle_row_of_nnz": (c_int32, [P, c_int32, c_int64]),
        "oracle_spmm_f32": (None, [c_int64, c_int64, P, P, P, P, c_int64, P, c_int64]),
        "oracle_spmm_f64": (None, [c_int64, c_int64, P, P, P, P, c_int64, P, c_int64]),
        "oracle_spmm_abs": (None, [c_int64, c_int64, P, P, P, P, c_int64, P, c_int64]),
        "oracle_segment_coo_f32": (None, [P, c_int64, P, c_int64, c_int64, P, c_int64]),
        "oracle_segment_coo_f64": (None, [P, c_int64, P, c_int64, c_int64, P, c_int64, P]),
        "oracle_check_result": (c_int64, [c_int64, c_int64, P, P]),
        "oracle_transpose": (None, [c_int64, c_int64, P, P, P, P, P]),
        "oracle_sddmm_f64": (None, [c_int64, c_int64, P, P, P, c_int64, P, c_int64, P, P]),
        "oracle_generate_index": (c_int32, [c_int32, c_int32, c_int32, c_int32, c_double, c_uint32, P]),
        "oracle_spmm_f64d": (None, [c_int64, c_int64, P, P, P, P, c_int64, P, c_int64]),
        "oracle_sddmm_f64d": (None, [c_int64, c_int64, P, P, P, c_int64, P, c_int64, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


def _c(a, dtype):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)


def row_of_nnz(rowptr, e: int) -> int:
    rp = _c(rowptr, np.int32)
    return int(load().oracle_row_of_nnz(_p(rp), rp.size - 1, int(e)))


def spmm_f32(rowptr, colind, vals, B, C=None):
    """Reference CPU SpMM (fp32, i -> p -> j order). C is accumulated into (zeros if None)."""
    rp, ci, va = _c(rowptr, np.int32), _c(colind, np.int32), _c(vals, np.float32)
    B = _c(B, np.float32)
    M, N = rp.size - 1, B.shape[1]
    C = np.zeros((M, N), np.float32) if C is None else C
    assert C.dtype == np.float32 and C.flags.c_contiguous
    load().oracle_spmm_f32(M, N, _p(rp), _p(ci), _p(va), _p(B), N, _p(C), N)
    return C


def spmm_f64(rowptr, colind, vals, B):
    rp, ci, va = _c(rowptr, np.int32), _c(colind, np.int32), _c(vals, np.float32)
    B = _c(B, np.float32)
    M, N = rp.size - 1, B.shape[1]
    C = np.zeros((M, N), np.float64)
    D = np.zeros((M, N), np.float64)
    lib = load()
    lib.oracle_spmm_f64(M, N, _p(rp), _p(ci), _p(va), _p(B), N, _p(C), N)
    lib.oracle_spmm_abs(M, N, _p(rp), _p(ci), _p(va), _p(B), N, _p(D), N)
    return C, D


def spmm_f64d(rowptr, colind, vals, B):
    """float64 operands end to end (vals may be None = all ones)."""
    rp, ci, va = _c(rowptr, np.int32), _c(colind, np.int32), _c(vals, np.float64)
    B = _c(B, np.float64)
    M, N = rp.size - 1, B.shape[1]
    C = np.zeros((M, N), np.float64)
    load().oracle_spmm_f64d(M, N, _p(rp), _p(ci), _p(va), _p(B), N, _p(C), N)
    return C


def sddmm_f64d(rowptr, colind, X, Y):
    rp, ci = _c(rowptr, np.int32), _c(colind, np.int32)
    X, Y = _c(X, np.float64), _c(Y, np.float64)
    M, N = rp.size - 1, X.shape[1]
    out = np.zeros(ci.size, np.float64)
    load().oracle_sddmm_f64d(M, N, _p(rp), _p(ci), _p(X), N, _p(Y), N, _p(out))
    return out


def segment_coo_f32(src, index, dst_len):
    src, idx = _c(src, np.float32), _c(index, np.int32)
    E, N = src.shape
    dst = np.zeros((dst_len, N), np.float32)
    load().oracle_segment_coo_f32(_p(src), N, _p(idx), E, N, _p(dst), N)
    return dst


def segment_coo_f64(src, index, dst_len):
    src, idx = _c(src, np.float32), _c(index, np.int32)
    E, N = src.shape
    dst = np.zeros((dst_len, N), np.float64)
    absd = np.zeros((dst_len, N), np.float64)
    load().oracle_segment_coo_f64(_p(src), N, _p(idx), E, N, _p(dst), N, _p(absd))
    return dst, absd


def check_result(C, C_ref) -> int:
    C, R = _c(C, np.float32), _c(C_ref, np.float32)
    M, N = C.shape
    return int(load().oracle_check_result(M, N, _p(C), _p(R)))


def transpose(nrow, ncol, rowptr, colind):
    rp, ci = _c(rowptr, np.int32), _c(colind, np.int32)
    nnz = ci.size
    colptr = np.zeros(ncol + 1, np.int32)
    rowind = np.zeros(nnz, np.int32)
    perm = np.zeros(nnz, np.int32)
    load().oracle_transpose(nrow, ncol, _p(rp), _p(ci), _p(colptr), _p(rowind), _p(perm))
    return colptr, rowind, perm


def sddmm_f64(rowptr, colind, X, Y):
    rp, ci = _c(rowptr, np.int32), _c(colind, np.int32)
    X, Y = _c(X, np.float32), _c(Y, np.float32)
    M, N = rp.size - 1, X.shape[1]
    out = np.zeros(ci.size, np.float64)
    absd = np.zeros(ci.size, np.float64)
    load().oracle_sddmm_f64(M, N, _p(rp), _p(ci), _p(X), N, _p(Y), N, _p(out), _p(absd))
    return out, absd


def generate_index(range_: int, min_seg: int, max_seg: int, total: int, cv: float, seed: int = 1):
    out = np.zeros(total, np.int32)
    n = load().oracle_generate_index(range_, min_seg, max_seg, total, cv, seed, _p(out))
    return out, int(n)


def normalized_error(C, C64, D) -> float:
    """max |C - C64| / D over entries with D > 0; entries with D == 0 must be exactly 0
    (returns inf otherwise)."""
    C = np.asarray(C, np.float64)
    diff = np.abs(C - C64)
    zero = D == 0
    if np.any(diff[zero] != 0):
        return float("inf")
    if not np.any(~zero):
        return 0.0
    return float(np.max(diff[~zero] / D[~zero]))
