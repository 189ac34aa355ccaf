"""The fwd + bwd step with learnable values kept in the transpose's order (two-pass
permutation into A's order, forward SpMM with carries, fused backward with dvals in the
transpose's order), captured as one hipGraph: every replay, with eager work and
synchronisations between replays, equals the eager step bit for bit (integer-valued operands:
exact whatever order the carries' atomic adds take). The same for the step whose forward reads
the values through the permutation plan (segscan_spmm_csr_permuted)."""
import numpy as np
import pytest
import torch

from segscan_amd import ops, synth

pytestmark = pytest.mark.gpu


def test_fwd_bwd_step_graph_replays_equal_eager(gpu):
    m = synth.powerlaw_csr(M=20000, K=16000, max_deg=3000, target_nnz=300_000, seed=41, values="int")
    A = m.to_device(gpu)
    N = 64
    rng = np.random.default_rng(42)
    B = torch.from_numpy(rng.integers(-2, 3, size=(m.ncol, N)).astype(np.float32)).to(gpu)
    dC = torch.from_numpy(rng.integers(-2, 3, size=(m.nrow, N)).astype(np.float32)).to(gpu)
    T = ops.csr_transpose(A)
    iperm = ops.invert_permutation(T.perm)
    vals_t = A.vals.index_select(0, T.perm.long())
    plan = ops.PermutePlan(iperm)
    vals_a = torch.empty_like(vals_t)
    A_perm = A.with_vals(vals_a)
    knobs = dict(group_factor=5, coarsen_factor=2, thread_nz=8, write_policy=2)
    ws_f = ops.spmm_workspace(A, N, torch.float32, "ebsr", knobs)
    ws_b = ops.spmm_backward_workspace(T, N, torch.float32)
    C = torch.empty((m.nrow, N), device=gpu)
    dB = torch.empty_like(B)
    dvals = torch.empty(m.nnz, device=gpu)

    def step():
        plan(vals_t, out=vals_a)
        ops.spmm(A_perm, B, variant="ebsr", knobs=knobs, out=C, workspace=ws_f)
        ops.spmm_backward(T, None, dC, B, dB=dB, dvals=dvals, workspace=ws_b, vals_t=vals_t, dvals_at_order=True)

    step()
    torch.cuda.synchronize()
    ref = [t.clone() for t in (vals_a, C, dB, dvals)]
    assert torch.equal(ref[0], A.vals)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        step()
    for r in range(3):
        for t in (vals_a, C, dB, dvals):
            t.fill_(float("nan"))  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip((vals_a, C, dB, dvals), ref):
            assert torch.equal(a, b), r


@pytest.mark.parametrize("thread_nz", [8, 32])
def test_fused_values_step_graph_replays_equal_eager(gpu, thread_nz):
    """The same step through segscan_spmm_csr_permuted (the plan's first pass, then the SpMM
    reading values from the staged buckets; workgroup rounds of 2048 and 8192 entries),
    captured as one hipGraph: replays equal the eager step and the unfused step, bit for bit."""
    m = synth.powerlaw_csr(M=20000, K=16000, max_deg=3000, target_nnz=300_000, seed=43, values="int")
    A = m.to_device(gpu)
    N = 64
    rng = np.random.default_rng(44)
    B = torch.from_numpy(rng.integers(-2, 3, size=(m.ncol, N)).astype(np.float32)).to(gpu)
    dC = torch.from_numpy(rng.integers(-2, 3, size=(m.nrow, N)).astype(np.float32)).to(gpu)
    T = ops.csr_transpose(A)
    vals_t = A.vals.index_select(0, T.perm.long())
    plan = ops.PermutePlan(ops.invert_permutation(T.perm))
    knobs = dict(group_factor=5, coarsen_factor=2, thread_nz=thread_nz, thread_per_block=256, write_policy=2)
    ws_f = ops.spmm_workspace(A, N, torch.float32, "ebsr", knobs)
    ws_b = ops.spmm_backward_workspace(T, N, torch.float32)
    C = torch.empty((m.nrow, N), device=gpu)
    dB = torch.empty_like(B)
    dvals = torch.empty(m.nnz, device=gpu)

    def step():
        plan.spmm(A, vals_t, B, variant="ebsr", knobs=knobs, out=C, workspace=ws_f)
        ops.spmm_backward(T, None, dC, B, dB=dB, dvals=dvals, workspace=ws_b, vals_t=vals_t, dvals_at_order=True)

    step()
    torch.cuda.synchronize()
    ref = [t.clone() for t in (C, dB, dvals)]
    assert torch.equal(ref[0], ops.spmm(A, B, variant="ebsr", knobs=knobs))  # A.vals: the same values
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        step()
    for r in range(3):
        for t in (C, dB, dvals):
            t.fill_(float("nan"))  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip((C, dB, dvals), ref):
            assert torch.equal(a, b), r
