"""Backward pieces on the GPU: the CSR -> CSC transpose must equal the
oracle's counting-sort transpose (test-cg/include/dataloader/dataloader.hpp:123-144)
bit for bit, SDDMM matches the float64 oracle at the normalized 1e-6 bound, and
A^T . dC through the transposed matrix matches the oracle."""
import numpy as np
import pytest
import torch

from segscan_amd import ops, synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [0, 1])
def test_transpose_matches_oracle(gpu, oracle, seed):
    mats = [
        synth.small_random_csr(300, 200, 0.05, seed=seed, empty_rows=0.3),
        synth.powerlaw_csr(M=3000, K=2000, max_deg=2000, target_nnz=30000, seed=seed),
    ]
    for m in mats:
        A = m.to_device(gpu)
        T = ops.csr_transpose(A)
        colptr, rowind, perm = oracle.transpose(m.nrow, m.ncol, m.rowptr, m.colind)
        assert np.array_equal(T.t.rowptr.cpu().numpy(), colptr)
        assert np.array_equal(T.t.colind.cpu().numpy(), rowind)
        assert np.array_equal(T.perm.cpu().numpy(), perm)
        assert np.array_equal(T.t.vals.cpu().numpy(), m.vals[perm])


@pytest.mark.parametrize("K,with_vals", [(2**20 + 3, True), (2**20 + 3, False), (1, True), (130000, True)])
def test_transpose_radix_widths(gpu, oracle, K, with_vals):
    """Column counts needing one, two and three radix passes, tails of partial tiles,
    and matrices without values (the sort then carries ones)."""
    M = 5000
    m = synth.small_random_csr(M, K, min(0.9, 11.0 / K), seed=K % 97, empty_rows=0.2)
    A = ops.CsrMatrix.from_numpy(m.nrow, m.ncol, m.rowptr, m.colind, m.vals if with_vals else None, device=gpu)
    T = ops.csr_transpose(A)
    colptr, rowind, perm = oracle.transpose(m.nrow, m.ncol, m.rowptr, m.colind)
    assert m.nnz % 4096 != 0
    assert np.array_equal(T.t.rowptr.cpu().numpy(), colptr)
    assert np.array_equal(T.t.colind.cpu().numpy(), rowind)
    assert np.array_equal(T.perm.cpu().numpy(), perm)
    if with_vals:
        assert np.array_equal(T.t.vals.cpu().numpy(), m.vals[perm])


def test_transpose_dense_column(gpu, oracle):
    # every row references column 5: one column with 3000 entries
    M, K = 3000, 64
    rowptr = np.arange(0, 2 * M + 1, 2, dtype=np.int32)
    colind = np.tile(np.array([5, 9], np.int32), M)
    A = ops.CsrMatrix.from_numpy(M, K, rowptr, colind, np.arange(2 * M, dtype=np.float32), device=gpu)
    T = ops.csr_transpose(A)
    colptr, rowind, perm = oracle.transpose(M, K, rowptr, colind)
    assert np.array_equal(T.t.rowptr.cpu().numpy(), colptr)
    assert np.array_equal(T.t.colind.cpu().numpy(), rowind)
    assert np.array_equal(T.perm.cpu().numpy(), perm)


@pytest.mark.parametrize("N", [1, 3, 16, 64, 100])
def test_sddmm(gpu, oracle, N):
    m = synth.powerlaw_csr(M=2000, K=1500, max_deg=1500, target_nnz=20000, seed=3)
    rng = np.random.default_rng(N)
    X = rng.standard_normal((m.nrow, N)).astype(np.float32)
    Y = rng.standard_normal((m.ncol, N)).astype(np.float32)
    out = ops.sddmm(m.to_device(gpu), torch.from_numpy(X).to(gpu), torch.from_numpy(Y).to(gpu)).cpu().numpy()
    ref, absd = oracle.sddmm_f64(m.rowptr, m.colind, X, Y)
    assert oracle.normalized_error(out, ref, absd) <= 1e-6


def test_spmm_transpose_product(gpu, oracle):
    m = synth.powerlaw_csr(M=2500, K=1800, max_deg=1800, target_nnz=25000, seed=4, values="int")
    rng = np.random.default_rng(4)
    dC = rng.integers(-2, 3, size=(m.nrow, 32)).astype(np.float32)
    T = ops.csr_transpose(m.to_device(gpu))
    dB = ops.spmm(T.t, torch.from_numpy(dC).to(gpu), variant="auto").cpu().numpy()
    dense = np.zeros((m.nrow, m.ncol), np.float64)
    dense[m.row_ids(), m.colind] = m.vals
    assert np.array_equal(dB, (dense.T @ dC).astype(np.float32))
