"""Many uneven graphs in one launch.

The graphs' CSR matrices are concatenated block-diagonally; one element-balanced
SpMM launch (grid-strided, so its workgroups stay resident and walk the whole
batch) covers every graph, whatever its size or skew. Per-graph sum pooling is
one segment_coo over the node -> graph index (sorted by construction). The
forward, pooling and the full backward (A^T . dH and the SDDMM for the values in
one fused walk over the cached device transpose, gather for the pooling) are
recorded once into
a hipGraph (torch.cuda.CUDAGraph) and replayed; every buffer, workspace
included, is allocated before capture, and no step issues hipMemset (zero
fills are kernels).

The baseline it is compared with is the per-graph launch loop: one SpMM and one
pooling launch per graph.
"""
from __future__ import annotations

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


class BatchedGcnStep:
    """H = A.X ; P = segment_sum(H, graph) ; backward for a loss with gradient dP
    (and optionally dH): dH_total = dH + gather(dP); dX = A^T.dH_total;
    dvals = SDDMM(dH_total, X) on A's pattern."""

    def __init__(self, batch: GraphBatch, N: int, dtype=torch.float32, variant="auto", knobs=None,
                 variant_t="auto", knobs_t=None, seg_variant="auto", seg_knobs=None, fused: bool = True,
                 knobs_bwd=None):
        self.b = batch
        A = batch.A
        dev = A.rowptr.device
        self.N = N
        self.dtype = dtype
        self.variant, self.knobs = variant, knobs
        self.variant_t, self.knobs_t = variant_t, knobs_t
        self.seg_variant, self.seg_knobs = seg_variant, seg_knobs
        T = ops.csr_transpose(A)
        self.T = T
        self.At = T.t.measured().with_vals(None if A.vals is None else A.vals.index_select(0, T.perm.long()))
        self.H = torch.empty((A.nrow, N), dtype=torch.float32, device=dev)
        self.P = torch.empty((batch.n_graphs, N), dtype=torch.float32, device=dev)
        # dP gathered per node (+ dH): needed only with a dH of its own or the separate kernels
        # (nodes x N fp32), so it is allocated on first use (the eager run before a capture)
        self.dHt = None
        self.dX = torch.empty((A.ncol, N), dtype=torch.float32, device=dev)
        self.dvals = torch.empty(max(A.nnz, 1), dtype=torch.float32, device=dev)[: A.nnz]
        self.ws_f = ops.spmm_workspace(A, N, dtype, variant, knobs)
        self.ws_t = ops.spmm_workspace(self.At, N, torch.float32, variant_t, knobs_t)
        # fused backward (dX and dvals from one walk over the transpose) where N fits one group of lanes
        self.fused, self.ws_b = False, None
        self.knobs_bwd = knobs_bwd  # fused backward shape (e.g. from autotune.tune_backward); None = built-in
        if fused:
            try:
                self.ws_b = ops.spmm_backward_workspace(T, N, torch.float32, knobs_bwd)
                self.fused = True
            except _abi.SegscanError as exc:
                if exc.status != _abi.EUNSUPPORTED:
                    raise
        info = ops.segment_describe(A.nrow, N, batch.n_graphs, torch.float32, seg_variant, seg_knobs)
        self.ws_s = torch.empty(max(1, info["workspace_bytes"]), dtype=torch.uint8, device=dev)
        self.graph = None

    def run(self, X: torch.Tensor, dP: torch.Tensor, dH: torch.Tensor | None = None) -> None:
        A = self.b.A
        if tuple(X.shape) != (A.ncol, self.N) or X.dtype != self.dtype:
            raise ValueError(f"X must be {self.dtype} of shape ({A.ncol}, {self.N})")
        if tuple(dP.shape) != (self.b.n_graphs, self.N) or dP.dtype != torch.float32:
            raise ValueError(f"dP must be float32 of shape ({self.b.n_graphs}, {self.N})")
        if dH is not None and (tuple(dH.shape) != (A.nrow, self.N) or dH.dtype != torch.float32):
            raise ValueError(f"dH must be float32 of shape ({A.nrow}, {self.N})")
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


class PerGraphLoop:
    """The baseline: one SpMM launch and one pooling launch per graph."""

    def __init__(self, csr: HostCsr, offsets: np.ndarray, N: int, device="cuda", variant="plain_ebpr", knobs=None,
                 tuner=None):
        self.N = N
        self.subs = []
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
            ws = ops.spmm_workspace(A, N, torch.float32, v, k)
            self.subs.append((r0, r1, A, v, k, ws))
        self.zero_index = {}

    def run(self, X: torch.Tensor, H: torch.Tensor, P: torch.Tensor) -> None:
        for g, (r0, r1, A, v, k, ws) in enumerate(self.subs):
            ops.spmm(A, X[r0:r1], v, k, out=H[r0:r1], workspace=ws)
            n = r1 - r0
            idx = self.zero_index.get(n)
            if idx is None:
                idx = torch.zeros(n, dtype=torch.int32, device=X.device)
                self.zero_index[n] = idx
            ops.segment_coo(H[r0:r1], idx, 1, "segment_coo", out=P[g:g + 1])
