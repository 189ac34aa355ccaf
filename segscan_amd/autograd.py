"""Differentiable SpMM and segment pooling — a test-layer torch.autograd.Function
over the same C ABI as the forward (libsegscan_hip.so); nothing here runs on the
CPU.

SpMM  C = A(vals) . B, A fixed pattern (rowptr, colind), vals and B learnable
(vals in CSR order, or with values_order="transpose" in the transpose's order: the
forward reads them through a two-pass permutation plan, whose second pass the SpMM takes
over where its launch allows (segscan_spmm_csr_permuted), or, with
forward_values="vals_map", through the inverse permutation inside the SpMM)
and the fused backward reads them and writes their gradient coalesced, with no random
access per entry; for CSR-order values the backward permutes them into the transpose's
order and the gradient back with two such plans, or reads and writes them through perm
inside the walk; backward_values="auto", the default, times both once per width):
    dB    = A^T . dC     CSR -> CSC transpose on the device (cached per pattern)
    dvals = SDDMM(dC, B) on A's pattern: dvals[p] = <dC[row(p)], B[colind[p]]>
    When dvals is wanted, both come from one fused walk over the transpose (one
    gather of each dC row per entry) if N fits one group of lanes (N <= 512);
    wider N runs the SpMM kernels on the transposed matrix and the SDDMM kernel
    separately. dB alone (values without gradient) is the SpMM on the transpose
    with the values permuted once and cached.
segment_coo sum / mean pooling  dst[s] = sum (or mean) of src rows with index s:
    dsrc  = gather(ddst)[index]  (divided by the segment count for mean)
"""
from __future__ import annotations

import torch

from . import _abi, ops


class SpmmPattern:
    """A CSR pattern with its device transpose computed once."""

    def __init__(self, A: ops.CsrMatrix, variant="auto", knobs=None, variant_t="auto", knobs_t=None,
                 fused: bool = True, knobs_bwd=None, values_order: str = "csr", forward_values: str = "permute",
                 backward_values: str = "auto"):
        if values_order not in ("csr", "transpose"):
            raise ValueError("values_order must be 'csr' or 'transpose'")
        if forward_values not in ("permute", "vals_map"):
            raise ValueError("forward_values must be 'permute' or 'vals_map'")
        if backward_values not in ("auto", "permute", "perm"):
            raise ValueError("backward_values must be 'auto', 'permute' or 'perm'")
        self.A = A
        self.values_order = values_order
        T = ops.csr_transpose(A)
        self.T = ops.CscMatrix(T.t.measured(), T.perm)  # longest column guides the "auto" plan of A^T
        self.variant, self.knobs = variant, knobs
        self.variant_t, self.knobs_t = variant_t, knobs_t
        self.fused = fused  # False: always the separate transposed SpMM + SDDMM kernels
        self.knobs_bwd = knobs_bwd  # fused backward shape (e.g. from autotune.tune_backward); None = built-in
        self._fused = {}
        self._vals_t, self._vals_t_key = None, None
        # values kept in the transpose's order: the forward reads entry p's value at iperm[p],
        # either from a copy permuted into A's order (two streaming passes) or inside the SpMM
        self.iperm = ops.invert_permutation(T.perm) if values_order == "transpose" else None
        self.forward_values = forward_values
        self.permute = (ops.PermutePlan(self.iperm) if values_order == "transpose" and forward_values == "permute"
                        and A.nnz > 0 else None)
        # values in A's order: the fused backward either takes them into the transpose's order
        # and its dvals back with two permutation plans (streaming passes, "permute") or reads
        # and writes them through perm inside the walk (one random 4-byte read and write per
        # entry, "perm"). Which is faster depends on the matrix (power-law 262144^2, N = 64:
        # 0.399 against 0.429 ms per fwd + bwd; the GCN graph, 131072 nodes: 0.267 against
        # 0.240 ms), so "auto" times both on the first backward of each width and B type and
        # keeps the faster. Both run the same walk over the same values: dvals are equal bit
        # for bit, and dB differs at most as two runs of either do (the order in which the
        # carries of rows that span chunks are added). Plans are built on first use.
        self.backward_values = backward_values
        self._to_t = self._to_csr = None
        self.backward_choice = {}  # (N, B dtype, dB wanted) -> "permute" | "perm" ("auto")

    def csr_permute_plans(self):
        """(to the transpose's order, back to A's order) permutation plans, or None."""
        if self.values_order != "csr" or self.backward_values == "perm" or self.A.nnz == 0:
            return None
        if self._to_t is None:
            self._to_t = ops.PermutePlan(self.T.perm)  # x_t[q] = x[perm[q]]
            self._to_csr = ops.PermutePlan(ops.invert_permutation(self.T.perm))  # x[p] = x_t[iperm[p]]
        return self._to_t, self._to_csr

    def csr_backward_mode(self, key, via_plans, in_walk) -> str:
        """The backward route for values in A's order: the fixed choice, or for "auto" the
        faster of the two on their first use for `key` (each run once to warm up, then three
        times between HIP events on the current stream). Under hipGraph capture nothing is
        timed and "permute" runs."""
        if self.backward_values != "auto":
            return self.backward_values
        if self.csr_permute_plans() is None:
            return "perm"
        if key in self.backward_choice:
            return self.backward_choice[key]
        if torch.cuda.is_current_stream_capturing():
            return "permute"
        ms = {}
        for mode, fn in (("permute", via_plans), ("perm", in_walk)):
            fn()
            start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            start.record()
            for _ in range(3):
                fn()
            stop.record()
            stop.synchronize()
            ms[mode] = start.elapsed_time(stop)
        self.backward_choice[key] = min(ms, key=ms.get)
        return self.backward_choice[key]

    def fused_ok(self, N: int) -> bool:
        """Whether the fused backward covers N columns (one group of lanes per dense row)."""
        if not self.fused:
            return False
        if N not in self._fused:
            try:
                ops.spmm_backward_workspace(self.T, N, knobs=self.knobs_bwd)
                self._fused[N] = True
            except _abi.SegscanError as exc:
                if exc.status != _abi.EUNSUPPORTED:
                    raise
                self._fused[N] = False
        return self._fused[N]

    def with_vals(self, vals: torch.Tensor) -> ops.CsrMatrix:
        """A with vals in this pattern's value order."""
        if self.values_order == "transpose":
            if self.permute is not None:
                return self.A.with_vals(self.permute(vals))
            return self.A.with_vals_map(vals, self.iperm)
        return self.A.with_vals(vals)

    def to_pattern_order(self, vals_csr: torch.Tensor) -> torch.Tensor:
        """CSR-order values in this pattern's value order (set-up of a parameter)."""
        if self.values_order == "transpose":
            return vals_csr.index_select(0, self.T.perm.long())
        return vals_csr

    def vals_in_t_order(self, vals: torch.Tensor) -> torch.Tensor:
        """vals permuted into the transpose's order, cached while vals is unchanged
        (same storage and version), as for a fixed normalized adjacency."""
        key = (vals.data_ptr(), vals._version, vals.numel())
        if self._vals_t_key != key:
            self._vals_t = vals.index_select(0, self.T.perm.long())
            self._vals_t_key = key
        return self._vals_t

    def transposed_with_vals(self, vals: torch.Tensor) -> ops.CsrMatrix:
        if self.values_order == "transpose":
            return self.T.t.with_vals(vals)
        return self.T.t.with_vals(self.vals_in_t_order(vals))


class SpmmFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vals: torch.Tensor, B: torch.Tensor, pattern: SpmmPattern):
        v32 = vals.detach().float().contiguous()
        if pattern.permute is not None:
            # values kept in the transpose's order, read through the plan (its second pass
            # fused into the SpMM where the launch allows)
            C = pattern.permute.spmm(pattern.A, v32, B.detach().contiguous(), variant=pattern.variant,
                                     knobs=pattern.knobs)
        else:
            C = ops.spmm(pattern.with_vals(v32), B.detach().contiguous(), variant=pattern.variant,
                         knobs=pattern.knobs)
        ctx.save_for_backward(vals, B)
        ctx.pattern = pattern
        return C

    @staticmethod
    def backward(ctx, dC: torch.Tensor):
        vals, B = ctx.saved_tensors
        pattern = ctx.pattern
        dC = dC.contiguous().float()
        dvals = dB = None
        want_b, want_v = ctx.needs_input_grad[1], ctx.needs_input_grad[0]
        if want_v and pattern.fused_ok(dC.size(1)):
            # learnable values: both gradients from one walk over the transpose
            v32 = vals.detach().float().contiguous()
            # a 16-bit B is read as stored (fp32 dC, fp32 accumulation)
            Bs = B.detach().contiguous()
            if Bs.dtype not in (torch.bfloat16, torch.float16):
                Bs = Bs.float()
            if pattern.values_order == "transpose":
                # values and their gradient in the walk's own order: no perm at all
                dB, dvals = ops.spmm_backward(pattern.T, None, dC, Bs, want_dB=want_b, knobs=pattern.knobs_bwd,
                                              vals_t=v32, dvals_at_order=True)
            else:
                def via_plans():
                    # values taken into the walk's order and dvals back by two streaming passes
                    to_t, to_csr = pattern.csr_permute_plans()
                    gB, dvals_t = ops.spmm_backward(pattern.T, None, dC, Bs, want_dB=want_b,
                                                    knobs=pattern.knobs_bwd, vals_t=to_t(v32), dvals_at_order=True)
                    return gB, to_csr(dvals_t)

                def in_walk():
                    # values read and dvals written through the permutation inside the kernel
                    return ops.spmm_backward(pattern.T, v32, dC, Bs, want_dB=want_b, knobs=pattern.knobs_bwd)

                mode = pattern.csr_backward_mode((dC.size(1), Bs.dtype, want_b), via_plans, in_walk)
                dB, dvals = via_plans() if mode == "permute" else in_walk()
            return dvals.to(vals.dtype), (None if dB is None else dB.to(B.dtype)), None
        # dB alone (fixed values): the SpMM kernels on the transpose, values in its order (cached)
        if ctx.needs_input_grad[1]:
            At = pattern.transposed_with_vals(vals.detach().float())
            dB = ops.spmm(At, dC, variant=pattern.variant_t, knobs=pattern.knobs_t).to(B.dtype)
        if ctx.needs_input_grad[0]:
            dvals = ops.sddmm(pattern.A, dC, B.detach().float().contiguous())
            dvals = pattern.to_pattern_order(dvals).to(vals.dtype)
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
