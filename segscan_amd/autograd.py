"""Differentiable SpMM and segment pooling — a test-layer torch.autograd.Function
over the same C ABI as the forward (libsegscan_hip.so); nothing here runs on the
CPU.

SpMM  C = A(vals) . B, A fixed pattern (rowptr, colind), vals and B learnable:
    dB    = A^T . dC     CSR -> CSC transpose on the device (cached per pattern),
                         then the same SpMM kernels on the transposed matrix
    dvals = SDDMM(dC, B) on A's pattern: dvals[p] = <dC[row(p)], B[colind[p]]>
segment_coo sum / mean pooling  dst[s] = sum (or mean) of src rows with index s:
    dsrc  = gather(ddst)[index]  (divided by the segment count for mean)
"""
from __future__ import annotations

import torch

from . import ops


class SpmmPattern:
    """A CSR pattern with its device transpose computed once."""

    def __init__(self, A: ops.CsrMatrix, variant="auto", knobs=None, variant_t="auto", knobs_t=None):
        self.A = A
        T = ops.csr_transpose(A)
        self.T = ops.CscMatrix(T.t.measured(), T.perm)  # longest column guides the "auto" plan of A^T
        self.variant, self.knobs = variant, knobs
        self.variant_t, self.knobs_t = variant_t, knobs_t

    def with_vals(self, vals: torch.Tensor) -> ops.CsrMatrix:
        return self.A.with_vals(vals)

    def transposed_with_vals(self, vals: torch.Tensor) -> ops.CsrMatrix:
        return self.T.t.with_vals(vals.index_select(0, self.T.perm.long()))


class SpmmFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vals: torch.Tensor, B: torch.Tensor, pattern: SpmmPattern):
        A = pattern.with_vals(vals.detach().float().contiguous())
        C = ops.spmm(A, B.detach().contiguous(), variant=pattern.variant, knobs=pattern.knobs)
        ctx.save_for_backward(vals, B)
        ctx.pattern = pattern
        return C

    @staticmethod
    def backward(ctx, dC: torch.Tensor):
        vals, B = ctx.saved_tensors
        pattern = ctx.pattern
        dC = dC.contiguous().float()
        dvals = dB = None
        if ctx.needs_input_grad[1]:
            At = pattern.transposed_with_vals(vals.detach().float())
            dB = ops.spmm(At, dC, variant=pattern.variant_t, knobs=pattern.knobs_t).to(B.dtype)
        if ctx.needs_input_grad[0]:
            dvals = ops.sddmm(pattern.A, dC, B.detach().float().contiguous()).to(vals.dtype)
        return dvals, dB, None


def spmm(vals: torch.Tensor, B: torch.Tensor, pattern: SpmmPattern) -> torch.Tensor:
    return SpmmFunction.apply(vals, B, pattern)


class SegmentPoolFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, src: torch.Tensor, index: torch.Tensor, dst_len: int, mean: bool):
        out = ops.segment_coo(src.detach().contiguous(), index, dst_len)
        counts = None
        if mean:
            counts = ops.segment_counts(index, dst_len)
            out = out / counts.clamp(min=1).unsqueeze(1).to(out.dtype)
        ctx.save_for_backward(index, counts if counts is not None else torch.empty(0, device=index.device))
        ctx.mean = mean
        ctx.src_dtype = src.dtype
        return out

    @staticmethod
    def backward(ctx, ddst: torch.Tensor):
        index, counts = ctx.saved_tensors
        dsrc = ops.segment_gather(ddst.contiguous().float(), index, counts if ctx.mean else None)
        return dsrc.to(ctx.src_dtype), None, None, None


def segment_pool(src: torch.Tensor, index: torch.Tensor, dst_len: int, mean: bool = False) -> torch.Tensor:
    return SegmentPoolFunction.apply(src, index, dst_len, mean)
