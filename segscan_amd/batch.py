"""Many uneven graphs in one launch.

The graphs' CSR matrices are concatenated block-diagonally. The library's batch
plan (``BatchPlan``, segscan_batch_plan_*) runs the whole step in one persistent
kernel launch: H = A.X, the per-graph sum pooling P, and the backward for
dH = dP[graph] (dX = A^T.dH, which on a block-diagonal A is colsum(A)[j] *
dP[g(j)], and dvals = SDDMM(dH, X), computed from the X rows the forward
gathers). ``BatchedGcnStep`` records it into a hipGraph (torch.cuda.CUDAGraph)
and replays it; every buffer is allocated before capture and no step issues
hipMemset.

With a node gradient dH of its own the step runs kernel by kernel instead: the
element-balanced SpMM, a segment_coo pooling pass and the fused backward over the
cached device transpose. The baseline both are compared with is the per-graph
launch loop: one SpMM and one pooling launch per graph, with its variant and
knobs from the tuning table's bucket for that graph's size and skew.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _abi, ops
from .synth import HostCsr


@dataclass
class GraphBatch:
    A: ops.CsrMatrix          # block-diagonal CSR of all graphs
    offsets: np.ndarray       # node offsets, graph g owns [offsets[g], offsets[g+1])
    graph_index: torch.Tensor  # int32 node -> graph id (sorted)
    n_graphs: int


def make_batch(csr: HostCsr, offsets: np.ndarray, device="cuda") -> GraphBatch:
    gi = np.repeat(np.arange(offsets.size - 1, dtype=np.int32), np.diff(offsets).astype(np.int64))
    return GraphBatch(csr.to_device(device), offsets.astype(np.int64), torch.from_numpy(gi).to(device),
                      int(offsets.size - 1))


class BatchPlan:
    """The library's batch plan (segscan_batch_plan_*): forward SpMM, per-graph sum pooling
    and the backward for dH = dP[graph] in one persistent kernel launch per run. The plan
    borrows the batch's device arrays and holds the column sums of its values (recompute
    them with set_values when the values change)."""

    def __init__(self, batch: GraphBatch, N: int, dtype=torch.float32, knobs=None, stream=None):
        lib = _abi.load()
        self.b, self.N, self.dtype = batch, N, dtype
        A = batch.A
        self._desc = A.desc()  # keeps A's pointers; the tensors stay referenced through batch
        offs = np.ascontiguousarray(batch.offsets, dtype=np.int64)
        handle = ctypes.c_void_p()
        st = lib.segscan_batch_plan_create(ctypes.byref(self._desc), offs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                           batch.n_graphs, N, ops._dtype_code(torch.empty(0, dtype=dtype)),
                                           _abi.knobs_ptr(knobs), ops._stream(stream), ctypes.byref(handle))
        _abi.check(st, "batch_plan_create")
        self._h = handle

    def describe(self) -> dict:
        info = _abi.BatchInfo()
        _abi.check(_abi.load().segscan_batch_plan_describe(self._h, ctypes.byref(info)), "batch_plan_describe")
        return info.as_dict()

    def set_values(self, vals: torch.Tensor | None, stream=None) -> None:
        if vals is not None and (vals.dtype != torch.float32 or vals.numel() != self.b.A.nnz or not vals.is_contiguous()):
            raise ValueError("vals must be contiguous float32 with nnz elements")
        self.b.A = self.b.A.with_vals(vals)
        self._desc = self.b.A.desc()
        _abi.check(_abi.load().segscan_batch_plan_set_values(self._h, ops._ptr(vals), ops._stream(stream)),
                   "batch_plan_set_values")

    def run(self, X: torch.Tensor, dP: torch.Tensor | None, P: torch.Tensor, H: torch.Tensor | None = None,
            dX: torch.Tensor | None = None, dvals: torch.Tensor | None = None, stream=None) -> None:
        A, N = self.b.A, self.N
        ops._require_device(X, dP, P, H, dX, dvals)
        if tuple(X.shape) != (A.ncol, N) or X.dtype != self.dtype:
            raise ValueError(f"X must be {self.dtype} of shape ({A.ncol}, {N})")
        if dP is not None and (tuple(dP.shape) != (self.b.n_graphs, N) or dP.dtype != torch.float32):
            raise ValueError(f"dP must be float32 of shape ({self.b.n_graphs}, {N})")
        for name, t, shape in (("P", P, (self.b.n_graphs, N)), ("H", H, (A.nrow, N)), ("dX", dX, (A.ncol, N))):
            if t is not None and (tuple(t.shape) != shape or t.dtype != torch.float32):
                raise ValueError(f"{name} must be float32 of shape {shape}")
        if dvals is not None and (dvals.dtype != torch.float32 or dvals.numel() != A.nnz or not dvals.is_contiguous()):
            raise ValueError("dvals must be contiguous float32 with nnz elements")

        def ld(t, name):
            return 0 if t is None else ops._leading_dim(t, name)
        st = _abi.load().segscan_batch_plan_run(self._h, ops._ptr(X), ld(X, "X"), ops._ptr(dP), ld(dP, "dP"),
                                               ops._ptr(H), ld(H, "H"), ops._ptr(P), ld(P, "P"), ops._ptr(dX),
                                               ld(dX, "dX"), ops._ptr(dvals), ops._stream(stream))
        _abi.check(st, "batch_plan_run")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _abi.load().segscan_batch_plan_destroy(h)
            except Exception:
                pass
            self._h = None


class BatchedGcnStep:
    """H = A.X ; P = segment_sum(H, graph) ; backward for a loss with gradient dP
    (and optionally dH): dH_total = dH + gather(dP); dX = A^T.dH_total;
    dvals = SDDMM(dH_total, X) on A's pattern.

    Without a dH of its own, the step is one launch of the library's batch plan
    (single_launch=True, the default). With a dH, or single_launch=False, it runs
    kernel by kernel: the SpMM, a segment_coo pooling pass and the fused backward over
    the cached device transpose (dP read through the graph index when there is no
    dH), or with fused=False the transposed SpMM and the SDDMM as separate kernels."""

    def __init__(self, batch: GraphBatch, N: int, dtype=torch.float32, variant="auto", knobs=None,
                 variant_t="auto", knobs_t=None, seg_variant="auto", seg_knobs=None, fused: bool = True,
                 knobs_bwd=None, single_launch: bool = True, plan_knobs=None):
        self.b = batch
        A = batch.A
        dev = A.rowptr.device
        self.N = N
        self.dtype = dtype
        self.variant, self.knobs = variant, knobs
        self.variant_t, self.knobs_t = variant_t, knobs_t
        self.seg_variant, self.seg_knobs = seg_variant, seg_knobs
        self.fused_requested = fused
        self.knobs_bwd = knobs_bwd  # fused backward shape (e.g. from autotune.tune_backward); None = built-in
        self.H = torch.empty((A.nrow, N), dtype=torch.float32, device=dev)
        self.P = torch.empty((batch.n_graphs, N), dtype=torch.float32, device=dev)
        self.dX = torch.empty((A.ncol, N), dtype=torch.float32, device=dev)
        self.dvals = torch.empty(max(A.nnz, 1), dtype=torch.float32, device=dev)[: A.nnz]
        self.plan = BatchPlan(batch, N, dtype, plan_knobs) if single_launch else None
        self._sep_ready = False
        # dP gathered per node (+ dH): needed only with a dH of its own or the separate kernels
        # (nodes x N fp32), so it is allocated on first use (the eager run before a capture)
        self.dHt = None
        self.graph = None
        if not single_launch:
            self._separate_setup()

    def _separate_setup(self) -> None:
        """Resources of the kernel-by-kernel form (transpose, workspaces), built once."""
        if self._sep_ready:
            return
        A, N, dev = self.b.A, self.N, self.b.A.rowptr.device
        T = ops.csr_transpose(A)
        self.T = T
        self.At = T.t.measured().with_vals(None if A.vals is None else A.vals.index_select(0, T.perm.long()))
        self.ws_f = ops.spmm_workspace(A, N, self.dtype, self.variant, self.knobs)
        self.ws_t = ops.spmm_workspace(self.At, N, torch.float32, self.variant_t, self.knobs_t)
        # fused backward (dX and dvals from one walk over the transpose) where N fits one group of lanes
        self.fused, self.ws_b = False, None
        if self.fused_requested:
            try:
                self.ws_b = ops.spmm_backward_workspace(T, N, torch.float32, self.knobs_bwd)
                self.fused = True
            except _abi.SegscanError as exc:
                if exc.status != _abi.EUNSUPPORTED:
                    raise
        info = ops.segment_describe(A.nrow, N, self.b.n_graphs, torch.float32, self.seg_variant, self.seg_knobs)
        self.ws_s = torch.empty(max(1, info["workspace_bytes"]), dtype=torch.uint8, device=dev)
        self._sep_ready = True

    def run(self, X: torch.Tensor, dP: torch.Tensor, dH: torch.Tensor | None = None) -> None:
        A = self.b.A
        if tuple(X.shape) != (A.ncol, self.N) or X.dtype != self.dtype:
            raise ValueError(f"X must be {self.dtype} of shape ({A.ncol}, {self.N})")
        if tuple(dP.shape) != (self.b.n_graphs, self.N) or dP.dtype != torch.float32:
            raise ValueError(f"dP must be float32 of shape ({self.b.n_graphs}, {self.N})")
        if dH is not None and (tuple(dH.shape) != (A.nrow, self.N) or dH.dtype != torch.float32):
            raise ValueError(f"dH must be float32 of shape ({A.nrow}, {self.N})")
        if self.plan is not None and dH is None:
            self.plan.run(X, dP, self.P, H=self.H, dX=self.dX, dvals=self.dvals)
            return
        self._separate_setup()
        ops.spmm(A, X, self.variant, self.knobs, out=self.H, workspace=self.ws_f)
        ops.segment_coo(self.H, self.b.graph_index, self.b.n_graphs, self.seg_variant, self.seg_knobs, out=self.P,
                        workspace=self.ws_s)
        if self.fused and dH is None:
            # dH_total = dP[graph]: the fused backward reads dP through the graph index, so the
            # gathered copy (nodes x N) is never written
            ops.spmm_backward(self.T, None, dP, X, dB=self.dX, dvals=self.dvals, workspace=self.ws_b,
                              vals_t=self.At.vals, knobs=self.knobs_bwd, dc_map=self.b.graph_index)
            return
        if self.dHt is None:
            self.dHt = torch.empty((A.nrow, self.N), dtype=torch.float32, device=dP.device)
        ops.segment_gather(dP, self.b.graph_index, out=self.dHt)
        if dH is not None:
            self.dHt.add_(dH)
        if self.fused:
            # the batch's values are fixed: read them in the transpose's order (cached at set-up);
            # a 16-bit X is read as stored
            ops.spmm_backward(self.T, None, self.dHt, X, dB=self.dX, dvals=self.dvals, workspace=self.ws_b,
                              vals_t=self.At.vals, knobs=self.knobs_bwd)
            return
        Xf = X if X.dtype == torch.float32 else X.float()
        ops.spmm(self.At, self.dHt, self.variant_t, self.knobs_t, out=self.dX, workspace=self.ws_t)
        ops.sddmm(A, self.dHt, Xf, out=self.dvals)

    def capture(self, X: torch.Tensor, dP: torch.Tensor, dH: torch.Tensor | None = None) -> torch.cuda.CUDAGraph:
        """Records run() into a hipGraph (after one eager warm-up on a side stream)."""
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.run(X, dP, dH)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run(X, dP, dH)
        self.graph = g
        return g


def table_tuner(dtype, N: int, table: dict | None = None):
    """A PerGraphLoop tuner: each graph's variant and knobs from the tuning table's bucket
    for its (N, row skew, size), else the built-in AUTO rule."""
    from . import autotune
    tab = autotune.load_table() if table is None else table

    def pick(sub: HostCsr):
        stats = autotune.row_stats(sub.rowptr)
        hit = autotune.lookup(tab, "spmm", dtype, N, stats)
        return (hit[0], hit[1]) if hit else ("auto", None)
    return pick


class PerGraphLoop:
    """The baseline: one SpMM launch and one pooling launch per graph (the reference's
    plain launcher by default; tuner=table_tuner(...) picks each graph's variant from the
    tuning table, with the AUTO segment_coo for its pooling)."""

    def __init__(self, csr: HostCsr, offsets: np.ndarray, N: int, device="cuda", variant="plain_ebpr", knobs=None,
                 tuner=None, dtype=torch.float32, seg_variant=None):
        self.N = N
        self.subs = []
        self.seg_variant = seg_variant or ("segment_coo" if tuner is None else "auto")
        rp = csr.rowptr.astype(np.int64)
        for g in range(offsets.size - 1):
            r0, r1 = int(offsets[g]), int(offsets[g + 1])
            p0, p1 = int(rp[r0]), int(rp[r1])
            sub_rp = (rp[r0:r1 + 1] - p0).astype(np.int32)
            sub_ci = (csr.colind[p0:p1].astype(np.int64) - r0).astype(np.int32)
            sub_va = csr.vals[p0:p1]
            A = ops.CsrMatrix.from_numpy(r1 - r0, r1 - r0, sub_rp, sub_ci, sub_va, device=device)
            v, k = variant, knobs
            if tuner is not None:
                v, k = tuner(HostCsr(r1 - r0, r1 - r0, sub_rp, sub_ci, sub_va))
            ws = ops.spmm_workspace(A, N, dtype, v, k)
            info = ops.segment_describe(r1 - r0, N, 1, torch.float32, self.seg_variant)
            ws_s = torch.empty(max(1, info["workspace_bytes"]), dtype=torch.uint8, device=device)
            self.subs.append((r0, r1, A, v, k, ws, ws_s))
        self.zero_index = {}

    def run(self, X: torch.Tensor, H: torch.Tensor, P: torch.Tensor) -> None:
        for g, (r0, r1, A, v, k, ws, ws_s) in enumerate(self.subs):
            ops.spmm(A, X[r0:r1], v, k, out=H[r0:r1], workspace=ws)
            n = r1 - r0
            idx = self.zero_index.get(n)
            if idx is None:
                idx = torch.zeros(n, dtype=torch.int32, device=X.device)
                self.zero_index[n] = idx
            ops.segment_coo(H[r0:r1], idx, 1, self.seg_variant, out=P[g:g + 1], workspace=ws_s)
