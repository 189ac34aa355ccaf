"""Torch-facing wrappers over the C ABI: CSR SpMM, segment_coo and the backward
pieces. Device memory comes from torch; every call goes to libsegscan_hip.so on
torch's current HIP stream. There is no CPU path: CPU tensors are rejected.
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_size_t
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _abi

_DTYPES = {torch.float32: _abi.F32, torch.bfloat16: _abi.BF16, torch.float16: _abi.F16}


def _dtype_code(t: torch.Tensor) -> int:
    try:
        return _DTYPES[t.dtype]
    except KeyError:
        raise TypeError(f"dense operand dtype {t.dtype} not supported (float32, bfloat16, float16)") from None


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(stream=None):
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def _require_device(*ts):
    for t in ts:
        if t is not None and not t.is_cuda:
            raise ValueError("segscan_amd operates on GPU tensors only (no CPU fallback)")


def _leading_dim(t: torch.Tensor, name: str) -> int:
    if t.dim() != 2:
        raise ValueError(f"{name} must be 2-D")
    if t.stride(1) != 1 and t.size(1) > 1:
        raise ValueError(f"{name} must be row-major with unit column stride")
    ld = t.stride(0)
    if t.size(0) <= 1:
        ld = max(ld, t.size(1))
    return ld


@dataclass
class CsrMatrix:
    """CSR matrix on the device; mirrors SpMatCsrDescr_t
    (test-cg/include/dataloader/dataloader.hpp:146-168) with int32 indices and
    nullable values (None = all ones, as __guard_load_default_one)."""

    nrow: int
    ncol: int
    rowptr: torch.Tensor
    colind: torch.Tensor
    vals: Optional[torch.Tensor] = None
    max_row_nnz: int = 0  # longest row, 0 = unknown; only guides variant="auto"
    vals_map: Optional[torch.Tensor] = None  # int32 [nnz]: entry p's value is vals[vals_map[p]]

    @property
    def nnz(self) -> int:
        return int(self.colind.numel())

    @classmethod
    def from_numpy(cls, nrow, ncol, rowptr, colind, vals=None, device="cuda") -> "CsrMatrix":
        rp_h = np.asarray(rowptr)
        max_row = int(np.diff(rp_h.astype(np.int64)).max()) if rp_h.size > 1 else 0
        rp = torch.as_tensor(rp_h, dtype=torch.int32).to(device)
        ci = torch.as_tensor(colind, dtype=torch.int32).to(device)
        va = None if vals is None else torch.as_tensor(vals, dtype=torch.float32).to(device)
        return cls(int(nrow), int(ncol), rp.contiguous(), ci.contiguous(), None if va is None else va.contiguous(),
                   max_row)

    def with_vals(self, vals: Optional[torch.Tensor]) -> "CsrMatrix":
        """The same pattern with values in CSR order."""
        return CsrMatrix(self.nrow, self.ncol, self.rowptr, self.colind, vals, self.max_row_nnz)

    def with_vals_map(self, vals: torch.Tensor, vals_map: torch.Tensor) -> "CsrMatrix":
        """The same pattern with values kept in another order: entry p's value is
        vals[vals_map[p]] (e.g. the transpose's order, vals_map = its inverse perm)."""
        if vals_map.dtype != torch.int32 or vals_map.numel() != self.nnz or not vals_map.is_contiguous():
            raise ValueError("vals_map must be contiguous int32 with nnz entries")
        return CsrMatrix(self.nrow, self.ncol, self.rowptr, self.colind, vals, self.max_row_nnz, vals_map)

    def measured(self) -> "CsrMatrix":
        """The same matrix with max_row_nnz read from the device (one synchronisation)."""
        m = int(torch.diff(self.rowptr).max()) if self.nrow > 0 else 0
        return CsrMatrix(self.nrow, self.ncol, self.rowptr, self.colind, self.vals, m, self.vals_map)

    def desc(self) -> _abi.CsrDesc:
        d = _abi.CsrDesc()
        d.nrow, d.ncol, d.nnz = self.nrow, self.ncol, self.nnz
        d.rowptr = self.rowptr.data_ptr() if self.rowptr.numel() else None
        d.colind = self.colind.data_ptr() if self.colind.numel() else None
        d.vals = None if self.vals is None or self.vals.numel() == 0 else self.vals.data_ptr()
        d.max_row_nnz = self.max_row_nnz
        d.vals_map = None if self.vals_map is None or self.vals_map.numel() == 0 else self.vals_map.data_ptr()
        return d


# ---- SpMM --------------------------------------------------------------------------------

def spmm_describe(A: CsrMatrix, N: int, dtype=torch.float32, variant="auto", knobs=None) -> dict:
    lib = _abi.load()
    info = _abi.LaunchInfo()
    v = _abi.SPMM_VARIANTS[variant] if isinstance(variant, str) else int(variant)
    _abi.check(lib.segscan_spmm_describe(byref(A.desc()), N, _DTYPES[dtype], v, _abi.knobs_ptr(knobs), byref(info)),
               "spmm_describe")
    return info.as_dict()


def spmm_workspace(A: CsrMatrix, N: int, dtype=torch.float32, variant="auto", knobs=None, device=None) -> torch.Tensor:
    info = spmm_describe(A, N, dtype, variant, knobs)
    return torch.empty(max(1, info["workspace_bytes"]), dtype=torch.uint8, device=device or A.rowptr.device)


def spmm(A: CsrMatrix, B: torch.Tensor, variant="auto", knobs=None, out: Optional[torch.Tensor] = None,
         accumulate: bool = False, workspace: Optional[torch.Tensor] = None, stream=None) -> torch.Tensor:
    """C (=|+=) A @ B with B of dtype float32/bfloat16/float16 and fp32 C."""
    _require_device(A.rowptr, B)
    lib = _abi.load()
    N = B.size(1)
    if B.size(0) != A.ncol:
        raise ValueError(f"B has {B.size(0)} rows, A has {A.ncol} columns")
    if out is None:
        out = torch.empty((A.nrow, N), dtype=torch.float32, device=B.device)
    if out.dtype != torch.float32 or tuple(out.shape) != (A.nrow, N):
        raise ValueError("out must be float32 of shape (nrow, N)")
    v = _abi.SPMM_VARIANTS[variant] if isinstance(variant, str) else int(variant)
    kp = _abi.knobs_ptr(knobs)
    desc = A.desc()
    if workspace is None:
        ws_bytes = c_size_t(0)
        _abi.check(lib.segscan_spmm_workspace_size(byref(desc), N, _dtype_code(B), v, kp, byref(ws_bytes)),
                   "spmm_workspace_size")
        workspace = torch.empty(max(1, ws_bytes.value), dtype=torch.uint8, device=B.device)
    st = lib.segscan_spmm_csr(byref(desc), N, _ptr(B), _leading_dim(B, "B"), _dtype_code(B), _ptr(out),
                              _leading_dim(out, "C"), v, kp, _abi.ACCUMULATE if accumulate else 0,
                              _ptr(workspace), workspace.numel(), _stream(stream))
    _abi.check(st, "spmm")
    return out


# ---- segment_coo ---------------------------------------------------------------------------

def segment_describe(E: int, N: int, dst_len: int, dtype=torch.float32, variant="auto", knobs=None) -> dict:
    lib = _abi.load()
    info = _abi.LaunchInfo()
    v = _abi.SEG_VARIANTS[variant] if isinstance(variant, str) else int(variant)
    _abi.check(lib.segscan_segment_coo_describe(E, N, dst_len, _DTYPES[dtype], v, _abi.knobs_ptr(knobs), byref(info)),
               "segment_coo_describe")
    return info.as_dict()


def segment_coo(src: torch.Tensor, index: torch.Tensor, dst_len: int, variant="auto", knobs=None,
                out: Optional[torch.Tensor] = None, accumulate: bool = False,
                workspace: Optional[torch.Tensor] = None, stream=None) -> torch.Tensor:
    """dst[s] (=|+=) sum of src rows e with index[e] == s (index sorted)."""
    _require_device(src, index)
    lib = _abi.load()
    E, N = src.shape
    if index.dtype != torch.int32 or index.numel() != E:
        raise ValueError("index must be int32 with one entry per src row")
    if out is None:
        out = torch.empty((dst_len, N), dtype=torch.float32, device=src.device)
    v = _abi.SEG_VARIANTS[variant] if isinstance(variant, str) else int(variant)
    kp = _abi.knobs_ptr(knobs)
    if workspace is None:
        ws_bytes = c_size_t(0)
        _abi.check(lib.segscan_segment_coo_workspace_size(E, N, dst_len, _dtype_code(src), v, kp, byref(ws_bytes)),
                   "segment_coo_workspace_size")
        workspace = torch.empty(max(1, ws_bytes.value), dtype=torch.uint8, device=src.device)
    st = lib.segscan_segment_coo(_ptr(src), _leading_dim(src, "src"), _dtype_code(src), _ptr(index), E, N, _ptr(out),
                                 _leading_dim(out, "dst"), dst_len, v, kp, _abi.ACCUMULATE if accumulate else 0,
                                 _ptr(workspace), workspace.numel(), _stream(stream))
    _abi.check(st, "segment_coo")
    return out


# ---- backward pieces ----------------------------------------------------------------------

@dataclass
class CscMatrix:
    """Transpose of a CsrMatrix as a CSR matrix of A^T (ncol x nrow) plus perm:
    perm[q] = CSR position of transposed entry q."""

    t: CsrMatrix
    perm: torch.Tensor


def csr_transpose(A: CsrMatrix, stream=None) -> CscMatrix:
    _require_device(A.rowptr)
    lib = _abi.load()
    dev = A.rowptr.device
    colptr = torch.empty(A.ncol + 1, dtype=torch.int32, device=dev)
    rowind = torch.empty(max(A.nnz, 1), dtype=torch.int32, device=dev)
    perm = torch.empty(max(A.nnz, 1), dtype=torch.int32, device=dev)
    vals_t = torch.empty(max(A.nnz, 1), dtype=torch.float32, device=dev)
    desc = A.desc()
    ws_bytes = c_size_t(0)
    _abi.check(lib.segscan_csr_transpose_workspace_size(byref(desc), byref(ws_bytes)), "transpose_workspace_size")
    ws = torch.empty(max(1, ws_bytes.value), dtype=torch.uint8, device=dev)
    _abi.check(lib.segscan_csr_transpose(byref(desc), _ptr(colptr), _ptr(rowind), _ptr(perm), _ptr(vals_t), _ptr(ws),
                                         ws.numel(), _stream(stream)), "csr_transpose")
    n = A.nnz
    t = CsrMatrix(A.ncol, A.nrow, colptr, rowind[:n], None if A.vals is None else vals_t[:n])
    return CscMatrix(t, perm[:n])


def sddmm(A: CsrMatrix, X: torch.Tensor, Y: torch.Tensor, out: Optional[torch.Tensor] = None, stream=None):
    """out[p] = <X[row(p)], Y[colind[p]]> on A's pattern (fp32 result)."""
    _require_device(X, Y)
    lib = _abi.load()
    if X.dtype != Y.dtype:
        raise TypeError("X and Y must share a dtype")
    N = X.size(1)
    if out is None:
        out = torch.empty(max(A.nnz, 1), dtype=torch.float32, device=X.device)[: A.nnz]
    _abi.check(lib.segscan_sddmm_csr(byref(A.desc()), N, _ptr(X), _leading_dim(X, "X"), _ptr(Y),
                                     _leading_dim(Y, "Y"), _dtype_code(X), _ptr(out), _stream(stream)), "sddmm")
    return out


def spmm_backward_workspace(T: CscMatrix, N: int, dtype=torch.float32, knobs=None, device=None) -> torch.Tensor:
    lib = _abi.load()
    ws_bytes = c_size_t(0)
    _abi.check(lib.segscan_spmm_backward_workspace_size(byref(T.t.desc()), N, _DTYPES[dtype], _abi.knobs_ptr(knobs),
                                                        byref(ws_bytes)), "spmm_backward_workspace_size")
    return torch.empty(max(1, ws_bytes.value), dtype=torch.uint8, device=device or T.t.rowptr.device)


def spmm_backward(T: CscMatrix, vals: Optional[torch.Tensor], dC: torch.Tensor, B: torch.Tensor, want_dB: bool = True,
                  want_dvals: bool = True, dB: Optional[torch.Tensor] = None, dvals: Optional[torch.Tensor] = None,
                  knobs=None, accumulate: bool = False, workspace: Optional[torch.Tensor] = None, stream=None,
                  vals_t: Optional[torch.Tensor] = None, dc_map: Optional[torch.Tensor] = None,
                  dvals_at_order: bool = False):
    """Fused backward of C = A @ B over T = csr_transpose(A), one gather of dC per entry:
    dB = A^T @ dC (K x N, fp32) and dvals[p] = <dC[row(p)], B[colind[p]]> (nnz, fp32, A's order).
    vals: A's values in A's order (None = ones), read through T.perm; vals_t: the same values
    in the transpose's order (e.g. T.t.vals while A's values are unchanged), read coalesced
    and used instead of vals. dc_map: int32, one entry per row of A; row r of A then reads
    dC[dc_map[r]] (dC may have any number of rows), e.g. dC = dP with dc_map = the pooling
    index, for the backward of a per-segment sum pooling without the gathered copy.
    dvals_at_order: dvals in the transpose's order (coalesced stores), the gradient of
    values kept in that order (with vals_t, the call reads no perm).
    Returns (dB or None, dvals or None)."""
    _require_device(dC, B)
    lib = _abi.load()
    if dC.dtype != B.dtype and dC.dtype != torch.float32:
        raise TypeError("dC and B must share a dtype, or dC be float32")
    K, M, nnz = T.t.nrow, T.t.ncol, T.t.nnz
    N = B.size(1)
    if dc_map is not None:
        _require_device(dc_map)
        if dc_map.dtype != torch.int32 or dc_map.numel() != M or not dc_map.is_contiguous():
            raise ValueError(f"dc_map must be contiguous int32 with {M} entries")
        if tuple(B.shape) != (K, N) or dC.dim() != 2 or dC.size(1) != N or (M > 0 and dC.size(0) < 1):
            raise ValueError(f"B must be ({K}, N) and dC (rows >= 1, N)")
    elif tuple(B.shape) != (K, N) or tuple(dC.shape) != (M, N):
        raise ValueError(f"B must be ({K}, N) and dC ({M}, N)")
    if want_dB and dB is None:
        dB = torch.empty((K, N), dtype=torch.float32, device=B.device)
    if want_dvals and dvals is None:
        dvals = torch.empty(max(nnz, 1), dtype=torch.float32, device=B.device)[:nnz]
    if dB is not None and (dB.dtype != torch.float32 or tuple(dB.shape) != (K, N)):
        raise ValueError("dB must be float32 of shape (K, N)")
    if dvals is not None and (dvals.dtype != torch.float32 or dvals.numel() != nnz):
        raise ValueError("dvals must be float32 with nnz elements")
    for name, v in (("vals", vals), ("vals_t", vals_t)):
        if v is not None and (v.dtype != torch.float32 or v.numel() != nnz or not v.is_contiguous()):
            raise ValueError(f"{name} must be contiguous float32 with nnz elements")
    kp = _abi.knobs_ptr(knobs)
    if workspace is None:
        workspace = spmm_backward_workspace(T, N, dC.dtype, knobs, B.device) if dB is not None else None
    desc = T.t.with_vals(vals_t).desc()  # the At-order values, if any, travel in At.vals
    tail = (_ptr(B), _leading_dim(B, "B"), _dtype_code(B), _ptr(dB), _leading_dim(dB, "dB") if dB is not None else 0,
            _ptr(dvals), kp, (_abi.ACCUMULATE if accumulate else 0) | (_abi.DVALS_AT_ORDER if dvals_at_order else 0),
            _ptr(workspace),
            0 if workspace is None else workspace.numel(), _stream(stream))
    if dc_map is not None:
        st = lib.segscan_spmm_backward_csr_rowmap(byref(desc), _ptr(T.perm), _ptr(vals), N, _ptr(dC),
                                                  _leading_dim(dC, "dC"), _dtype_code(dC), _ptr(dc_map), *tail)
    else:
        st = lib.segscan_spmm_backward_csr(byref(desc), _ptr(T.perm), _ptr(vals), N, _ptr(dC), _leading_dim(dC, "dC"),
                                           _dtype_code(dC), *tail)
    _abi.check(st, "spmm_backward")
    return dB, dvals


def invert_permutation(perm: torch.Tensor, stream=None) -> torch.Tensor:
    """inv[perm[i]] = i (int32, on the device): the vals_map that reads values kept in the
    transpose's order."""
    _require_device(perm)
    if perm.dtype != torch.int32 or not perm.is_contiguous():
        raise ValueError("perm must be contiguous int32")
    inv = torch.empty_like(perm)
    _abi.check(_abi.load().segscan_invert_permutation(_ptr(perm), perm.numel(), _ptr(inv), _stream(stream)),
               "invert_permutation")
    return inv


class PermutePlan:
    """out[i] = x[src[i]] for a fixed permutation src of [0, n) (segscan_permute_plan_*): two
    streaming passes over LDS-staged tiles instead of one random 4-byte read per value. For
    values kept in the transpose's order, src = invert_permutation(T.perm) gives them in A's
    order. Calls of one plan share its staging buffer: order them on one stream."""

    def __init__(self, src: torch.Tensor, stream=None):
        _require_device(src)
        if src.dtype != torch.int32 or src.dim() != 1 or not src.is_contiguous() or src.numel() < 1:
            raise ValueError("src must be a non-empty contiguous 1-D int32 tensor")
        self.n = src.numel()
        self.device = src.device
        handle = ctypes.c_void_p()
        _abi.check(_abi.load().segscan_permute_plan_create(_ptr(src), self.n, _stream(stream), ctypes.byref(handle)),
                   "permute_plan_create")
        self._h = handle

    def __call__(self, x: torch.Tensor, out: Optional[torch.Tensor] = None, stream=None) -> torch.Tensor:
        _require_device(x)
        if x.dtype != torch.float32 or x.numel() != self.n or not x.is_contiguous():
            raise ValueError(f"x must be contiguous float32 with {self.n} elements")
        if out is None:
            out = torch.empty_like(x)
        elif out.dtype != torch.float32 or out.numel() != self.n or not out.is_contiguous():
            raise ValueError(f"out must be contiguous float32 with {self.n} elements")
        _abi.check(_abi.load().segscan_permute_run(self._h, _ptr(x), _ptr(out), _stream(stream)), "permute_run")
        return out

    def spmm(self, A: "CsrMatrix", vals_src: torch.Tensor, B: torch.Tensor, variant="auto", knobs=None,
             out: Optional[torch.Tensor] = None, accumulate: bool = False,
             workspace: Optional[torch.Tensor] = None, stream=None) -> torch.Tensor:
        """C (=|+=) A'.B with A' = A's pattern holding value p = vals_src[src[p]]
        (segscan_spmm_csr_permuted): equal bit for bit to spmm(A.with_vals(self(vals_src)), B, ...).
        An ebsr launch whose workgroup rounds cover part of one 8192-value bucket reads the
        values from the plan's staging buffer (the second pass fused into the SpMM); other
        launches run both passes. A's own values are ignored. The workspace is spmm's for the
        same A, N, dtype, variant and knobs."""
        _require_device(A.rowptr, B, vals_src)
        if A.nnz != self.n:
            raise ValueError(f"plan permutes {self.n} values, A has {A.nnz} entries")
        if vals_src.dtype != torch.float32 or vals_src.numel() != self.n or not vals_src.is_contiguous():
            raise ValueError(f"vals_src must be contiguous float32 with {self.n} elements")
        lib = _abi.load()
        N = B.size(1)
        if B.size(0) != A.ncol:
            raise ValueError(f"B has {B.size(0)} rows, A has {A.ncol} columns")
        if out is None:
            out = torch.empty((A.nrow, N), dtype=torch.float32, device=B.device)
        if out.dtype != torch.float32 or tuple(out.shape) != (A.nrow, N):
            raise ValueError("out must be float32 of shape (nrow, N)")
        v = _abi.SPMM_VARIANTS[variant] if isinstance(variant, str) else int(variant)
        kp = _abi.knobs_ptr(knobs)
        desc = A.desc()
        desc.vals = None
        desc.vals_map = None
        if workspace is None:
            ws_bytes = c_size_t(0)
            _abi.check(lib.segscan_spmm_workspace_size(byref(desc), N, _dtype_code(B), v, kp, byref(ws_bytes)),
                       "spmm_workspace_size")
            workspace = torch.empty(max(1, ws_bytes.value), dtype=torch.uint8, device=B.device)
        st = lib.segscan_spmm_csr_permuted(byref(desc), self._h, _ptr(vals_src), N, _ptr(B), _leading_dim(B, "B"),
                                           _dtype_code(B), _ptr(out), _leading_dim(out, "C"), v, kp,
                                           _abi.ACCUMULATE if accumulate else 0, _ptr(workspace), workspace.numel(),
                                           _stream(stream))
        _abi.check(st, "spmm_csr_permuted")
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _abi.load().segscan_permute_plan_destroy(h)
            except Exception:
                pass
            self._h = None


def segment_counts(index: torch.Tensor, dst_len: int, stream=None) -> torch.Tensor:
    _require_device(index)
    lib = _abi.load()
    counts = torch.empty(max(dst_len, 1), dtype=torch.int32, device=index.device)[:dst_len]
    _abi.check(lib.segscan_segment_counts(_ptr(index), index.numel(), dst_len, _ptr(counts), _stream(stream)),
               "segment_counts")
    return counts


def segment_gather(ddst: torch.Tensor, index: torch.Tensor, counts: Optional[torch.Tensor] = None,
                   out: Optional[torch.Tensor] = None, stream=None) -> torch.Tensor:
    """dsrc[e] = ddst[index[e]] (divided by the segment count when counts is given)."""
    _require_device(ddst, index)
    lib = _abi.load()
    E = index.numel()
    N = ddst.size(1)
    if out is None:
        out = torch.empty((E, N), dtype=torch.float32, device=ddst.device)
    _abi.check(lib.segscan_segment_gather(_ptr(ddst), _leading_dim(ddst, "ddst"), _ptr(index), E, N, _ptr(counts),
                                          _ptr(out), _leading_dim(out, "dsrc"), _stream(stream)), "segment_gather")
    return out


def check_csr(A: CsrMatrix, stream=None) -> int:
    """Device-side structure check; returns the flag bits (0 = well formed)."""
    lib = _abi.load()
    flags = torch.zeros(1, dtype=torch.int32, device=A.rowptr.device)
    _abi.check(lib.segscan_check_csr(byref(A.desc()), _ptr(flags), _stream(stream)), "check_csr")
    return int(flags.item())


def check_index(index: torch.Tensor, dst_len: int, stream=None) -> int:
    lib = _abi.load()
    flags = torch.zeros(1, dtype=torch.int32, device=index.device)
    _abi.check(lib.segscan_check_index(_ptr(index), index.numel(), dst_len, _ptr(flags), _stream(stream)),
               "check_index")
    return int(flags.item())
