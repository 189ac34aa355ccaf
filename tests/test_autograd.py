"""Backward of SpMM and segment pooling.

CPU (float64): the backward formulas used on the GPU — dB = A^T.dC through the
counting-sort transpose, dvals = SDDMM(dC, B) on A's pattern, and the gather
(divided by counts for mean pooling) — restated on the oracle in float64 and
checked by torch.autograd.gradcheck against central finite differences
(eps = 1e-6) on a 512-node GCN-normalized subgraph.

GPU: the test-layer autograd Function over the HIP library (fp32) against the
same float64 gradients, normalized error <= 1e-6.
"""
import numpy as np
import pytest
import torch

from segscan_amd import synth


@pytest.fixture(scope="module")
def subgraph():
    return synth.gcn_normalized(n=512, avg_deg=20, seed=3)


class OracleSpmm(torch.autograd.Function):
    """float64 SpMM with the GPU's backward formulas, computed by the oracle."""

    @staticmethod
    def forward(ctx, vals, B, m, tr, oracle):
        ctx.save_for_backward(vals, B)
        ctx.m, ctx.tr, ctx.oracle = m, tr, oracle
        return torch.from_numpy(oracle.spmm_f64d(m.rowptr, m.colind, vals.detach().numpy(), B.detach().numpy()))

    @staticmethod
    def backward(ctx, dC):
        vals, B = ctx.saved_tensors
        m, (colptr, rowind, perm), o = ctx.m, ctx.tr, ctx.oracle
        dC = dC.detach().numpy()
        dB = o.spmm_f64d(colptr, rowind, vals.detach().numpy()[perm], dC)
        dvals = o.sddmm_f64d(m.rowptr, m.colind, dC, B.detach().numpy())
        return torch.from_numpy(dvals), torch.from_numpy(dB), None, None, None


def _gather_mean_f64(src, index, dst_len, mean):
    """Segment pooling in float64 with the GPU's backward (gather, / count)."""

    class F(torch.autograd.Function):
        @staticmethod
        def forward(ctx, s):
            out = torch.zeros((dst_len, s.shape[1]), dtype=torch.float64).index_add_(0, index.long(), s)
            counts = torch.bincount(index.long(), minlength=dst_len).clamp(min=1).to(torch.float64)
            ctx.counts = counts
            return out / counts[:, None] if mean else out

        @staticmethod
        def backward(ctx, d):
            g = d[index.long()]
            return g / ctx.counts[index.long()][:, None] if mean else g

    return F.apply(src)


def test_spmm_backward_formulas_match_finite_differences(oracle, subgraph):
    m = subgraph
    tr = oracle.transpose(m.nrow, m.ncol, m.rowptr, m.colind)
    rng = np.random.default_rng(0)
    vals = torch.from_numpy(m.vals.astype(np.float64)).requires_grad_()
    B = torch.from_numpy(rng.standard_normal((m.ncol, 128))).requires_grad_()
    assert torch.autograd.gradcheck(lambda v, b: OracleSpmm.apply(v, b, m, tr, oracle), (vals, B), eps=1e-6,
                                    atol=1e-6, rtol=1e-6, fast_mode=True)


@pytest.mark.parametrize("mean", [False, True])
def test_segment_pool_backward_matches_finite_differences(mean):
    rng = np.random.default_rng(1)
    deg = np.clip(rng.zipf(2.1, size=60), 0, 40)
    deg[::7] = 0
    index = torch.from_numpy(synth.segment_index_from_degrees(deg))
    src = torch.from_numpy(rng.standard_normal((index.numel(), 16))).requires_grad_()
    assert torch.autograd.gradcheck(lambda s: _gather_mean_f64(s, index, deg.size, mean), (src,), eps=1e-6,
                                    atol=1e-6, rtol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_gpu_spmm_gradients_match_fp64(gpu, oracle, subgraph, dtype):
    from segscan_amd import autograd as ag
    m = subgraph
    rng = np.random.default_rng(2)
    A = m.to_device(gpu)
    pattern = ag.SpmmPattern(A)
    B32 = rng.standard_normal((m.ncol, 128)).astype(np.float32)
    B = torch.from_numpy(B32).to(dtype).to(gpu).requires_grad_()
    vals = torch.from_numpy(m.vals).to(gpu).requires_grad_()
    C = ag.spmm(vals, B, pattern)
    dC = rng.standard_normal(C.shape).astype(np.float32)
    C.backward(torch.from_numpy(dC).to(gpu))
    Bq = B.detach().float().cpu().numpy()  # the operand values the GPU read
    # forward
    C64, D = oracle.spmm_f64(m.rowptr, m.colind, m.vals, Bq)
    assert oracle.normalized_error(C.detach().cpu().numpy(), C64, D) <= 1e-6
    # dB = A^T dC
    colptr, rowind, perm = oracle.transpose(m.nrow, m.ncol, m.rowptr, m.colind)
    dB64 = oracle.spmm_f64d(colptr, rowind, m.vals.astype(np.float64)[perm], dC.astype(np.float64))
    dBabs = oracle.spmm_f64d(colptr, rowind, np.abs(m.vals.astype(np.float64))[perm], np.abs(dC.astype(np.float64)))
    dB_gpu = B.grad.float().cpu().numpy()
    tol = 1e-6 if dtype == torch.float32 else 8e-3  # bf16 gradient storage rounds to 2^-9
    assert oracle.normalized_error(dB_gpu, dB64, dBabs) <= tol
    # dvals = SDDMM(dC, B)
    dv64 = oracle.sddmm_f64d(m.rowptr, m.colind, dC.astype(np.float64), Bq.astype(np.float64))
    dvabs = oracle.sddmm_f64d(m.rowptr, m.colind, np.abs(dC.astype(np.float64)), np.abs(Bq.astype(np.float64)))
    assert oracle.normalized_error(vals.grad.cpu().numpy(), dv64, dvabs) <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("mean", [False, True])
def test_gpu_segment_pool_gradients(gpu, mean):
    from segscan_amd import autograd as ag
    rng = np.random.default_rng(3)
    deg = np.clip(rng.zipf(2.1, size=500), 0, 300)
    deg[::9] = 0
    idx = synth.segment_index_from_degrees(deg)
    src = torch.from_numpy(rng.standard_normal((idx.size, 64)).astype(np.float32)).to(gpu).requires_grad_()
    index = torch.from_numpy(idx).to(gpu)
    out = ag.segment_pool(src, index, deg.size, mean=mean)
    ref = torch.zeros((deg.size, 64), dtype=torch.float64).index_add_(0, torch.from_numpy(idx).long(),
                                                                      src.detach().cpu().double())
    counts = torch.from_numpy(np.maximum(deg, 1)).double()
    if mean:
        ref = ref / counts[:, None]
    np.testing.assert_allclose(out.detach().cpu().numpy(), ref.numpy(), rtol=1e-5, atol=1e-5)
    d = rng.standard_normal((deg.size, 64)).astype(np.float32)
    out.backward(torch.from_numpy(d).to(gpu))
    expect = d[idx] / (counts.numpy()[idx][:, None] if mean else 1.0)
    np.testing.assert_allclose(src.grad.cpu().numpy(), expect, rtol=1e-6, atol=1e-7)


def test_backward_candidates_cover_n():
    """Every candidate shape of the fused-backward tuner covers N with at most 64 lanes,
    and the built-in shape comes first."""
    from segscan_amd import autotune
    for N in (1, 7, 64, 128, 200, 512):
        cands = autotune.backward_candidates(N)
        assert cands[0] is None
        for k in cands[1:]:
            assert -(-N // k["coarsen_factor"]) <= 64 and k["thread_nz"] in (8, 16)
    assert len(autotune.backward_candidates(128)) == 7  # 2, 4 and 8 columns per lane (64, 32, 16 lanes)
    assert len(autotune.backward_candidates(512)) == 3  # only 8 columns per lane fit in 64 lanes


@pytest.mark.gpu
def test_gpu_backward_tuner_and_tuned_shape(gpu, oracle, subgraph):
    """tune_backward times every candidate shape (fastest first); the autograd layer
    with the tuned shape gives the same float64-checked gradients."""
    from segscan_amd import autograd as ag, autotune
    m = subgraph
    rng = np.random.default_rng(4)
    A = m.to_device(gpu)
    N = 128
    B32 = rng.standard_normal((m.ncol, N)).astype(np.float32)
    dC = rng.standard_normal((m.nrow, N)).astype(np.float32)
    pattern = ag.SpmmPattern(A)
    rows = autotune.tune_backward(pattern.T, A.vals, torch.from_numpy(dC).to(gpu), torch.from_numpy(B32).to(gpu),
                                  autotune.backward_candidates(N), reps=2)
    ms = [r["ms"] for r in rows]
    assert ms == sorted(ms) and np.isfinite(ms[0]) and len(rows) == len(autotune.backward_candidates(N))
    for knobs in (rows[0]["knobs"], {"coarsen_factor": 8, "thread_nz": 8}):
        pattern = ag.SpmmPattern(A, knobs_bwd=knobs)
        B = torch.from_numpy(B32).to(gpu).requires_grad_()
        vals = torch.from_numpy(m.vals).to(gpu).requires_grad_()
        ag.spmm(vals, B, pattern).backward(torch.from_numpy(dC).to(gpu))
        colptr, rowind, perm = oracle.transpose(m.nrow, m.ncol, m.rowptr, m.colind)
        dB64 = oracle.spmm_f64d(colptr, rowind, m.vals.astype(np.float64)[perm], dC.astype(np.float64))
        dBabs = oracle.spmm_f64d(colptr, rowind, np.abs(m.vals.astype(np.float64))[perm], np.abs(dC.astype(np.float64)))
        assert oracle.normalized_error(B.grad.cpu().numpy(), dB64, dBabs) <= 1e-6
        dv64 = oracle.sddmm_f64d(m.rowptr, m.colind, dC.astype(np.float64), B32.astype(np.float64))
        dvabs = oracle.sddmm_f64d(m.rowptr, m.colind, np.abs(dC.astype(np.float64)), np.abs(B32.astype(np.float64)))
        assert oracle.normalized_error(vals.grad.cpu().numpy(), dv64, dvabs) <= 1e-6
