"""hipSPARSE SpMM as a timing comparator (the reference compared against
cuSPARSE, test-cg/include/spmm/spmm.cuh:640-694 and README §5). Loaded from
libsegscan_vendor.so, separate from the product library."""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_int, c_int64, c_void_p
from pathlib import Path

import torch

from .ops import CsrMatrix

LIB = Path(__file__).resolve().parent / "libsegscan_vendor.so"
_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise ImportError(f"{LIB} is missing (built by segscan_amd.build)")
        lib = ctypes.CDLL(str(LIB))
        lib.vendor_spmm_create.restype = c_int
        lib.vendor_spmm_create.argtypes = [c_int64, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_int64,
                                           c_void_p, c_int, c_void_p, c_int, c_void_p, POINTER(c_void_p)]
        lib.vendor_spmm_run.restype = c_int
        lib.vendor_spmm_run.argtypes = [c_void_p]
        lib.vendor_spmm_destroy.restype = None
        lib.vendor_spmm_destroy.argtypes = [c_void_p]
        _lib = lib
    return _lib


class VendorSpmm:
    """C = A @ B through hipsparseSpMM (row-major dense, fp32 operands and compute;
    the library has no fp32-A / 16-bit-B form)."""

    ALGS = {"default": 0, "csr_alg1": 1, "csr_alg2": 2, "csr_alg3": 3}

    def __init__(self, A: CsrMatrix, B: torch.Tensor, C: torch.Tensor, alg: str = "default"):
        lib = load()
        if A.vals is None:
            raise ValueError("hipSPARSE needs explicit values")
        if B.dtype != torch.float32 or C.dtype != torch.float32:
            raise ValueError("the hipSPARSE comparator takes fp32 B and C only")
        dt = 0
        self._keep = (A, B, C)
        h = c_void_p()
        st = lib.vendor_spmm_create(A.nrow, A.ncol, A.nnz, A.rowptr.data_ptr(), A.colind.data_ptr(),
                                    A.vals.data_ptr(), B.size(1), B.data_ptr(), dt, C.data_ptr(), self.ALGS[alg],
                                    c_void_p(torch.cuda.current_stream().cuda_stream), ctypes.byref(h))
        if st != 0:
            raise RuntimeError(f"hipSPARSE SpMM setup failed with status {st}")
        self._h = h

    def __call__(self):
        st = load().vendor_spmm_run(self._h)
        if st != 0:
            raise RuntimeError(f"hipsparseSpMM failed with status {st}")

    def __del__(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.vendor_spmm_destroy(self._h)
            self._h = None
