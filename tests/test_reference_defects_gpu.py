"""Regression tests for defects of the reference that this build does not
reproduce:

* padding lanes adding 0 * B[k] (k = an out-of-tile nnz position) into row M-1
  or the last segment (test-cg/include/spmm/spmm.cuh:57-60,
  new-cg/include/segscan/segscan.cuh:36-37), which turns a non-finite operand
  value into NaN where the product has none;
* 32-bit row offsets (row * ldC, k * ldB) that wrap once M * N >= 2^31.

Non-finite checks are exact: values are nonzero integers, so every finite sum
is exact in any order, and +inf / NaN propagate the same way in any order.
"""
import numpy as np
import pytest
import torch

from segscan_amd import ops, synth

pytestmark = pytest.mark.gpu

SPMM_VARIANTS = ["plain_ebpr", "plain_rbpr", "plain_ebsr", "plain_rbsr", "ebpr", "rbpr", "ebsr", "rbsr", "auto"]
SEG_VARIANTS = ["segment_coo", "segment_coo_new", "segment_coo_sr", "segment_coo_noshmem_sr",
                "segment_coo_noshmem_lessatom_sr", "segment_coo_sorted_sr", "auto"]


def _nonfinite_case():
    rng = np.random.default_rng(12)
    M, K, N = 1500, 400, 40
    m = synth.small_random_csr(M, K, 0.03, seed=12, empty_rows=0.2)
    # column 0 is never referenced; its B row is NaN
    colind = m.colind.copy()
    colind[colind == 0] = 1
    keep = np.ones(colind.size, bool)
    for r in range(M):  # drop duplicates the remap may have made
        a, b = m.rowptr[r], m.rowptr[r + 1]
        _, first = np.unique(colind[a:b], return_index=True)
        mask = np.zeros(b - a, bool)
        mask[first] = True
        keep[a:b] = mask
    deg = np.array([keep[m.rowptr[r]:m.rowptr[r + 1]].sum() for r in range(M)])
    rowptr = np.concatenate(([0], np.cumsum(deg))).astype(np.int32)
    colind = colind[keep].astype(np.int32)
    vals = rng.choice(np.array([-2, -1, 1, 2], np.float32), size=colind.size)
    B = rng.integers(-2, 3, size=(K, N)).astype(np.float32)
    B[0, :] = np.nan
    inf_cols = {int(colind[0]), int(colind[-1])}
    for c in inf_cols:
        B[c, 3] = np.inf
    vals[np.isin(colind, list(inf_cols))] = np.abs(vals[np.isin(colind, list(inf_cols))])  # +inf only, no inf - inf
    return M, K, N, rowptr, colind, vals, B


@pytest.mark.parametrize("variant", SPMM_VARIANTS)
def test_spmm_nonfinite_operand_values_do_not_leak(gpu, oracle, variant):
    M, K, N, rowptr, colind, vals, B = _nonfinite_case()
    A = ops.CsrMatrix.from_numpy(M, K, rowptr, colind, vals, device=gpu)
    C = ops.spmm(A, torch.from_numpy(B).to(gpu), variant=variant).cpu().numpy()
    ref, _ = oracle.spmm_f64(rowptr, colind, vals, B)
    assert not np.isnan(ref).any()  # the NaN row of B is never referenced
    np.testing.assert_array_equal(C, ref.astype(np.float32))


@pytest.mark.parametrize("variant", SEG_VARIANTS)
def test_segment_nonfinite_src_does_not_leak(gpu, oracle, variant):
    rng = np.random.default_rng(13)
    deg = rng.integers(0, 9, size=300)
    deg[-1] = 3
    idx = synth.segment_index_from_degrees(deg)
    src = rng.choice(np.array([-2, -1, 1, 2], np.float32), size=(idx.size, 24))
    src[0, 5] = np.inf
    src[-1, 7] = np.inf
    ref = oracle.segment_coo_f64(src.astype(np.float64), idx, deg.size)[0]
    out = ops.segment_coo(torch.from_numpy(src).to(gpu), torch.from_numpy(idx).to(gpu), deg.size,
                          variant=variant).cpu().numpy()
    assert not np.isnan(ref).any()
    np.testing.assert_array_equal(out, ref.astype(np.float32))


@pytest.mark.parametrize("variant", ["auto", "ebsr", "rbpr", "plain_ebpr"])
def test_spmm_offsets_beyond_int32(gpu, variant):
    """M * N = 2.18e9 elements of C: row r holds vals[r] * B[r % K]."""
    M, K, N = 2**21 + 3, 64, 1040
    assert M * N > 2**31
    r = torch.arange(M, dtype=torch.int32, device=gpu)
    A = ops.CsrMatrix(M, K, torch.arange(M + 1, dtype=torch.int32, device=gpu), r % K,
                      ((r % 3) + 1).float(), max_row_nnz=1)
    g = torch.Generator(device="cpu").manual_seed(1)
    B = torch.randint(-2, 3, (K, N), generator=g).float().to(gpu)
    C = ops.spmm(A, B, variant=variant)
    check = torch.cat([torch.arange(64), torch.arange(M - 128, M),
                       torch.randint(0, M, (256,), generator=g)]).to(gpu)
    got = C.index_select(0, check)
    want = B.index_select(0, (check % K)) * ((check % 3) + 1).float()[:, None]
    assert torch.equal(got, want)
    del C


def test_segment_offsets_beyond_int32(gpu):
    """E * N = 2.18e9 elements of src: dst[s] = src[2s] + src[2s+1]."""
    E, N = 2**21 + 3, 1040
    dst_len = (E + 1) // 2
    g = torch.Generator(device=gpu).manual_seed(2)
    src = torch.randint(-2, 3, (E, N), generator=g, device=gpu, dtype=torch.float32)
    idx = (torch.arange(E, dtype=torch.int32, device=gpu) // 2)
    dst = ops.segment_coo(src, idx, dst_len, variant="auto")
    for s in (0, 1, dst_len // 2, dst_len - 2):
        assert torch.equal(dst[s], src[2 * s] + src[2 * s + 1])
    assert torch.equal(dst[dst_len - 1], src[E - 1])  # E is odd: the last segment has one row
    del src, dst
