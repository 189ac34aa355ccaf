"""Batched uneven graphs: the block-diagonal launch (forward, per-graph sum
pooling, full backward) against the oracle, its hipGraph replays against the
eager result (several replays with eager work and synchronisation between
them), and the per-graph launch loop against the same result."""
import numpy as np
import pytest
import torch

from segscan_amd import batch, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def graphs():
    return synth.batched_graphs(n_graphs=24, min_nodes=16, max_nodes=3000, max_deg=400, seed=5)


def _reference(oracle, csr, offsets, X, dP):
    H = oracle.spmm_f32(csr.rowptr, csr.colind, csr.vals, X)
    gi = np.repeat(np.arange(offsets.size - 1), np.diff(offsets))
    P = oracle.segment_coo_f32(H, gi.astype(np.int32), offsets.size - 1)
    dHt = dP[gi]
    colptr, rowind, perm = oracle.transpose(csr.nrow, csr.ncol, csr.rowptr, csr.colind)
    dX64, dXabs = oracle.spmm_f64(colptr, rowind, csr.vals[perm], dHt)
    dv64, dvabs = oracle.sddmm_f64(csr.rowptr, csr.colind, dHt, X)
    return H, P, (dX64, dXabs), (dv64, dvabs)


def _check(oracle, step, ref):
    H, P, (dX64, dXabs), (dv64, dvabs) = ref
    np.testing.assert_allclose(step.H.cpu().numpy(), H, rtol=1e-5, atol=1e-4)
    np.testing.assert_allclose(step.P.cpu().numpy(), P, rtol=1e-4, atol=1e-3)
    assert oracle.normalized_error(step.dX.cpu().numpy(), dX64, dXabs) <= 1e-6
    assert oracle.normalized_error(step.dvals.cpu().numpy(), dv64, dvabs) <= 1e-6


def test_batched_step_and_graph_replays(gpu, oracle, graphs):
    csr, offsets = graphs
    b = batch.make_batch(csr, offsets, gpu)
    rng = np.random.default_rng(0)
    N = 64
    step = batch.BatchedGcnStep(b, N)
    X = torch.empty((csr.nrow, N), device=gpu)
    dP = torch.empty((b.n_graphs, N), device=gpu)
    Xh = rng.uniform(-1, 1, size=(csr.nrow, N)).astype(np.float32)
    dPh = rng.standard_normal((b.n_graphs, N)).astype(np.float32)
    X.copy_(torch.from_numpy(Xh))
    dP.copy_(torch.from_numpy(dPh))
    step.run(X, dP)
    torch.cuda.synchronize()
    ref = _reference(oracle, csr, offsets, Xh, dPh)
    _check(oracle, step, ref)
    g = step.capture(X, dP)
    for rep in range(3):
        # eager work between replays: new inputs written in place, outputs scribbled over
        Xh = rng.uniform(-1, 1, size=(csr.nrow, N)).astype(np.float32)
        dPh = rng.standard_normal((b.n_graphs, N)).astype(np.float32)
        X.copy_(torch.from_numpy(Xh))
        dP.copy_(torch.from_numpy(dPh))
        step.H.fill_(123.0)
        step.dX.fill_(-7.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        _check(oracle, step, _reference(oracle, csr, offsets, Xh, dPh))


def test_per_graph_loop_matches(gpu, oracle, graphs):
    csr, offsets = graphs
    rng = np.random.default_rng(1)
    N = 32
    loop = batch.PerGraphLoop(csr, offsets, N, device=gpu, variant="plain_ebpr")
    Xh = rng.uniform(-1, 1, size=(csr.nrow, N)).astype(np.float32)
    X = torch.from_numpy(Xh).to(gpu)
    H = torch.empty((csr.nrow, N), device=gpu)
    P = torch.zeros((offsets.size - 1, N), device=gpu)
    loop.run(X, H, P)
    Href = oracle.spmm_f32(csr.rowptr, csr.colind, csr.vals, Xh)
    np.testing.assert_allclose(H.cpu().numpy(), Href, rtol=1e-5, atol=1e-4)
    gi = np.repeat(np.arange(offsets.size - 1), np.diff(offsets)).astype(np.int32)
    np.testing.assert_allclose(P.cpu().numpy(), oracle.segment_coo_f32(Href, gi, offsets.size - 1), rtol=1e-4,
                               atol=1e-3)


def test_batched_bf16_features(gpu, oracle, graphs):
    """bf16 X (the configuration of the 256-graph measurement): H, dX and dvals against
    the float64 oracle on the rounded X, eager and after capture + replays."""
    csr, offsets = graphs
    b = batch.make_batch(csr, offsets, gpu)
    rng = np.random.default_rng(2)
    N = 64
    X = torch.from_numpy(rng.uniform(-1, 1, size=(csr.nrow, N)).astype(np.float32)).to(torch.bfloat16).to(gpu)
    dPh = rng.standard_normal((b.n_graphs, N)).astype(np.float32)
    dP = torch.from_numpy(dPh).to(gpu)
    step = batch.BatchedGcnStep(b, N, dtype=torch.bfloat16)
    Xq = X.float().cpu().numpy()
    H64, D = oracle.spmm_f64(csr.rowptr, csr.colind, csr.vals, Xq)
    gi = np.repeat(np.arange(offsets.size - 1), np.diff(offsets))
    dHt = dPh[gi]
    colptr, rowind, perm = oracle.transpose(csr.nrow, csr.ncol, csr.rowptr, csr.colind)
    dX64, dXabs = oracle.spmm_f64(colptr, rowind, csr.vals[perm], dHt)
    dv64, dvabs = oracle.sddmm_f64(csr.rowptr, csr.colind, dHt, Xq)

    def check():
        assert oracle.normalized_error(step.H.cpu().numpy(), H64, D) <= 1e-6
        assert oracle.normalized_error(step.dX.cpu().numpy(), dX64, dXabs) <= 1e-6
        assert oracle.normalized_error(step.dvals.cpu().numpy(), dv64, dvabs) <= 1e-6

    step.run(X, dP)
    torch.cuda.synchronize()
    check()
    g = step.capture(X, dP)
    for _ in range(2):
        step.H.fill_(9.0)
        step.dX.fill_(-7.0)
        step.dvals.fill_(3.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        check()


@pytest.mark.parametrize("fused", [True, False])
def test_batched_step_with_extra_dH(gpu, oracle, graphs, fused):
    """With a node gradient dH of its own the step forms dP[graph] + dH explicitly (the
    row-map backward does not apply); fused=False runs the transposed SpMM and SDDMM
    as separate kernels. Both against the oracle, eager and replayed."""
    csr, offsets = graphs
    b = batch.make_batch(csr, offsets, gpu)
    rng = np.random.default_rng(7)
    N = 32
    step = batch.BatchedGcnStep(b, N, fused=fused)
    Xh = rng.uniform(-1, 1, size=(csr.nrow, N)).astype(np.float32)
    dPh = rng.standard_normal((b.n_graphs, N)).astype(np.float32)
    dHh = rng.standard_normal((csr.nrow, N)).astype(np.float32)
    X, dP, dH = (torch.from_numpy(a).to(gpu) for a in (Xh, dPh, dHh))
    gi = np.repeat(np.arange(offsets.size - 1), np.diff(offsets))
    dtot = dPh[gi] + dHh
    colptr, rowind, perm = oracle.transpose(csr.nrow, csr.ncol, csr.rowptr, csr.colind)
    dX64, dXabs = oracle.spmm_f64(colptr, rowind, csr.vals[perm], dtot)
    dv64, dvabs = oracle.sddmm_f64(csr.rowptr, csr.colind, dtot, Xh)
    g = step.capture(X, dP, dH)
    for run in (lambda: step.run(X, dP, dH), g.replay):
        step.dX.fill_(-7.0)
        run()
        torch.cuda.synchronize()
        assert oracle.normalized_error(step.dX.cpu().numpy(), dX64, dXabs) <= 1e-6
        assert oracle.normalized_error(step.dvals.cpu().numpy(), dv64, dvabs) <= 1e-6
