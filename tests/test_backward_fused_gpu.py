"""Fused SpMM backward (segscan_spmm_backward_csr) on the GPU: one walk over the
transpose gives dB = A^T . dC and dvals = SDDMM(dC, B) on A's pattern.

Both outputs are checked against the float64 oracle (the transposed product through
the oracle's own transpose, the SDDMM through oracle_sddmm) at the normalized 1e-6
bound, and bit for bit where the operands are small integers (every partial sum is
then exact in fp32, in any order). The shapes cover skewed and even transposes,
empty rows and columns, carries across chunks (small thread_nz), partial column
groups (N not a multiple of the columns per lane), 16-bit operands, the optional
outputs and the accumulate flag.
"""
import numpy as np
import pytest
import torch

from segscan_amd import _abi, ops, synth

pytestmark = pytest.mark.gpu


def _oracle_grads(oracle, m, dC, B):
    """dB = A^T dC and dvals = <dC[row], B[col]> in float64, with their |.| bounds."""
    colptr, rowind, perm = oracle.transpose(m.nrow, m.ncol, m.rowptr, m.colind)
    vals_t = (m.vals if m.vals is not None else np.ones(m.nnz, np.float32))[perm]
    dB, dB_abs = oracle.spmm_f64(colptr, rowind, vals_t, dC)
    dv, dv_abs = oracle.sddmm_f64(m.rowptr, m.colind, dC, B)
    return dB, dB_abs, dv, dv_abs


def _matrices():
    return {
        "powerlaw_rows": synth.powerlaw_csr(M=6000, K=5000, max_deg=3000, target_nnz=60000, seed=7),
        "symmetric_powerlaw": synth.gcn_normalized(n=4000, avg_deg=12.0, seed=3),
        "empty_rows_cols": synth.small_random_csr(900, 1200, 0.01, seed=5, empty_rows=0.4),
        "uniform": synth.poisson_csr(M=3000, K=3000, mean=9.0, seed=2),
    }


MATS = _matrices()


@pytest.mark.parametrize("name", sorted(MATS))
@pytest.mark.parametrize("N", [1, 5, 16, 64, 100])
def test_fused_backward_matches_oracle(gpu, oracle, name, N):
    m = MATS[name]
    rng = np.random.default_rng(N * 31 + len(name))
    dC = rng.standard_normal((m.nrow, N)).astype(np.float32)
    B = rng.standard_normal((m.ncol, N)).astype(np.float32)
    A = m.to_device(gpu)
    T = ops.csr_transpose(A)
    dB, dv = ops.spmm_backward(T, A.vals, torch.from_numpy(dC).to(gpu), torch.from_numpy(B).to(gpu))
    rB, rB_abs, rv, rv_abs = _oracle_grads(oracle, m, dC, B)
    assert oracle.normalized_error(dB.cpu().numpy(), rB, rB_abs) <= 1e-6
    assert oracle.normalized_error(dv.cpu().numpy(), rv, rv_abs) <= 1e-6


@pytest.mark.parametrize("knobs", [
    None,
    {"thread_nz": 1},                                      # chunk = one tile: carries everywhere
    {"thread_nz": 3, "thread_per_block": 64},
    {"group_factor": 6, "coarsen_factor": 1},              # one column per lane, 64 lanes
    {"block_numer": 1, "block_denom": 1000000},            # one workgroup: grid-stride over every chunk
])
def test_fused_backward_exact_integers(gpu, knobs):
    m0 = MATS["powerlaw_rows"]
    rng = np.random.default_rng(17)
    m = synth.HostCsr(m0.nrow, m0.ncol, m0.rowptr, m0.colind,
                      rng.integers(-3, 4, size=m0.nnz).astype(np.float32))
    N = 48
    dC = rng.integers(-2, 3, size=(m.nrow, N)).astype(np.float32)
    B = rng.integers(-2, 3, size=(m.ncol, N)).astype(np.float32)
    A = m.to_device(gpu)
    T = ops.csr_transpose(A)
    dB, dv = ops.spmm_backward(T, A.vals, torch.from_numpy(dC).to(gpu), torch.from_numpy(B).to(gpu), knobs=knobs)
    dense = np.zeros((m.nrow, m.ncol), np.float64)
    dense[m.row_ids(), m.colind] = m.vals
    assert np.array_equal(dB.cpu().numpy(), (dense.T @ dC.astype(np.float64)).astype(np.float32))
    want = np.einsum("pn,pn->p", dC[m.row_ids()].astype(np.float64), B[m.colind].astype(np.float64))
    assert np.array_equal(dv.cpu().numpy(), want.astype(np.float32))


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16])
def test_fused_backward_16bit_operands(gpu, oracle, dtype):
    """16-bit dC and B, fp32 accumulation: against the oracle on the rounded operands."""
    m = MATS["powerlaw_rows"]
    N = 64
    rng = np.random.default_rng(11)
    dC = torch.from_numpy(rng.standard_normal((m.nrow, N)).astype(np.float32)).to(dtype)
    B = torch.from_numpy(rng.standard_normal((m.ncol, N)).astype(np.float32)).to(dtype)
    A = m.to_device(gpu)
    T = ops.csr_transpose(A)
    dB, dv = ops.spmm_backward(T, A.vals, dC.to(gpu), B.to(gpu))
    rB, rB_abs, rv, rv_abs = _oracle_grads(oracle, m, dC.float().numpy(), B.float().numpy())
    assert oracle.normalized_error(dB.cpu().numpy(), rB, rB_abs) <= 1e-6
    assert oracle.normalized_error(dv.cpu().numpy(), rv, rv_abs) <= 1e-6


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16])
def test_fused_backward_fp32_grad_16bit_operand(gpu, oracle, dtype):
    """fp32 dC with B kept in 16-bit storage (read as stored): against the oracle on
    the rounded B, and bit-equal to the fused kernel on an fp32 copy of the rounded B."""
    m = MATS["symmetric_powerlaw"]
    N = 64
    rng = np.random.default_rng(21)
    dC = torch.from_numpy(rng.standard_normal((m.nrow, N)).astype(np.float32)).to(gpu)
    B16 = torch.from_numpy(rng.standard_normal((m.ncol, N)).astype(np.float32)).to(dtype).to(gpu)
    A = m.to_device(gpu)
    T = ops.csr_transpose(A)
    dB, dv = ops.spmm_backward(T, A.vals, dC, B16)
    rB, rB_abs, rv, rv_abs = _oracle_grads(oracle, m, dC.cpu().numpy(), B16.float().cpu().numpy())
    assert oracle.normalized_error(dB.cpu().numpy(), rB, rB_abs) <= 1e-6
    assert oracle.normalized_error(dv.cpu().numpy(), rv, rv_abs) <= 1e-6
    _, dv32 = ops.spmm_backward(T, A.vals, dC, B16.float(), want_dB=False)
    assert torch.equal(dv, dv32)  # same products and order: widening is exact


def test_fused_backward_optional_outputs_and_accumulate(gpu, oracle):
    m = MATS["symmetric_powerlaw"]
    N = 32
    rng = np.random.default_rng(5)
    dC = torch.from_numpy(rng.standard_normal((m.nrow, N)).astype(np.float32)).to(gpu)
    B = torch.from_numpy(rng.standard_normal((m.ncol, N)).astype(np.float32)).to(gpu)
    A = m.to_device(gpu)
    T = ops.csr_transpose(A)
    full_dB, full_dv = ops.spmm_backward(T, A.vals, dC, B)
    dB_only, none_v = ops.spmm_backward(T, A.vals, dC, B, want_dvals=False)
    none_B, dv_only = ops.spmm_backward(T, A.vals, dC, B, want_dB=False)
    assert none_v is None and none_B is None
    # carries of At rows that span several chunks are added atomically, in any order
    torch.testing.assert_close(dB_only, full_dB, rtol=0, atol=1e-5)
    assert torch.equal(dv_only, full_dv)  # dvals: one plain store per entry
    # values in the transpose's order (coalesced) take precedence over A-order values
    dB_t, dv_t = ops.spmm_backward(T, torch.zeros_like(A.vals), dC, B, vals_t=T.t.vals)
    torch.testing.assert_close(dB_t, full_dB, rtol=0, atol=1e-5)
    assert torch.equal(dv_t, full_dv)
    base = torch.from_numpy(rng.standard_normal((m.ncol, N)).astype(np.float32)).to(gpu)
    acc = base.clone()
    ops.spmm_backward(T, A.vals, dC, B, dB=acc, want_dvals=False, accumulate=True)
    torch.testing.assert_close(acc, base + full_dB, rtol=0, atol=1e-5)
    # pattern-only A: values are ones
    ones_dB, _ = ops.spmm_backward(T, None, dC, B, want_dvals=False)
    rB, rB_abs, _, _ = _oracle_grads(oracle, synth.HostCsr(m.nrow, m.ncol, m.rowptr, m.colind,
                                                           np.ones(m.nnz, np.float32)), dC.cpu().numpy(),
                                     B.cpu().numpy())
    assert oracle.normalized_error(ones_dB.cpu().numpy(), rB, rB_abs) <= 1e-6


def test_fused_backward_trailing_empty_columns_and_sizes(gpu):
    """Columns of A past the last referenced one must come out as zero rows of dB
    (the last chunk owns them); nnz = 0 and N = 0 are handled without a kernel."""
    M, K, N = 50, 400, 8
    rowptr = np.arange(0, M + 1, dtype=np.int32)
    colind = (np.arange(M, dtype=np.int32) * 3) % 100  # only columns < 100
    vals = np.arange(1, M + 1, dtype=np.float32)
    A = ops.CsrMatrix.from_numpy(M, K, rowptr, colind, vals, device=gpu)
    T = ops.csr_transpose(A)
    dC = torch.ones((M, N), device=gpu)
    B = torch.ones((K, N), device=gpu)
    dB = torch.full((K, N), 7.0, device=gpu)
    ops.spmm_backward(T, A.vals, dC, B, dB=dB, want_dvals=False)
    want = np.zeros((K, N), np.float32)
    np.add.at(want, colind, vals[:, None] * np.ones((M, N), np.float32))
    assert np.array_equal(dB.cpu().numpy(), want)
    # no entries: dB is zero-filled, dvals empty
    E = ops.CsrMatrix.from_numpy(3, 5, np.zeros(4, np.int32), np.zeros(0, np.int32), device=gpu)
    TE = ops.csr_transpose(E)
    dB0, dv0 = ops.spmm_backward(TE, None, torch.ones((3, 4), device=gpu), torch.ones((5, 4), device=gpu))
    assert torch.count_nonzero(dB0) == 0 and dv0.numel() == 0
    # N = 0
    dBz, dvz = ops.spmm_backward(T, A.vals, torch.ones((M, 0), device=gpu), torch.ones((K, 0), device=gpu))
    assert dBz.shape == (K, 0) and torch.count_nonzero(dvz) == 0


def test_fused_backward_rejects_too_wide_N(gpu):
    m = MATS["uniform"]
    A = m.to_device(gpu)
    T = ops.csr_transpose(A)
    N = 600  # wider than 64 lanes x 8 columns
    with pytest.raises(_abi.SegscanError) as exc:
        ops.spmm_backward(T, A.vals, torch.ones((m.nrow, N), device=gpu), torch.ones((m.ncol, N), device=gpu))
    assert exc.value.status == _abi.EUNSUPPORTED


def _short_row_matrix(seed):
    """At rows of ~3 entries (nnz < 8 K, the row-batch form's condition), with empty
    columns of A (empty At rows) between and at the end."""
    rng = np.random.default_rng(seed)
    m = synth.small_random_csr(3000, 4000, 0.0008, seed=seed)
    keep = np.ones(4000, bool)
    keep[rng.choice(4000, size=600, replace=False)] = False
    keep[-40:] = False
    mask = keep[m.colind]
    deg = np.bincount(m.row_ids()[mask], minlength=m.nrow)
    rowptr = np.zeros(m.nrow + 1, np.int32)
    np.cumsum(deg, out=rowptr[1:])
    colind = m.colind[mask].astype(np.int32)
    vals = rng.integers(-3, 4, size=colind.size).astype(np.float32)
    return synth.HostCsr(m.nrow, m.ncol, rowptr, colind, vals)


@pytest.mark.parametrize("dc_dtype", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("N,knobs", [
    (32, None),                                            # 8 lanes x 4 columns
    (64, None),                                            # 16 x 4
    (128, None),                                           # 32 x 4
    (64, {"thread_nz": 1}),                                # carries at nearly every chunk
    (64, {"block_numer": 1, "block_denom": 1000000}),      # one workgroup, grid-stride over all chunks
    (61, {"coarsen_factor": 4}),                           # a partial last lane (1 of 4 columns)
])
def test_fused_backward_short_rows_bf16_operand(gpu, dc_dtype, N, knobs):
    """Short At rows with a bf16 B take the row-batch form (B rows loaded with each
    batch of dC rows); integer operands make every product and sum exact."""
    m = _short_row_matrix(N + 7)
    assert m.nnz < 8 * m.ncol and m.nnz > 0
    rng = np.random.default_rng(N)
    dC = rng.integers(-3, 4, size=(m.nrow, N)).astype(np.float32)
    B = rng.integers(-3, 4, size=(m.ncol, N)).astype(np.float32)
    A = m.to_device(gpu)
    T = ops.csr_transpose(A)
    dB = torch.full((m.ncol, N), 5.0, device=gpu)  # every row, empty ones too, must be written
    _, dv = ops.spmm_backward(T, A.vals, torch.from_numpy(dC).to(gpu).to(dc_dtype),
                              torch.from_numpy(B).to(gpu).to(torch.bfloat16), dB=dB, knobs=knobs)
    rows = m.row_ids()
    want_dB = np.zeros((m.ncol, N))
    np.add.at(want_dB, m.colind, m.vals[:, None].astype(np.float64) * dC[rows])
    want_dv = np.einsum("pn,pn->p", dC[rows].astype(np.float64), B[m.colind].astype(np.float64))
    assert np.array_equal(dB.cpu().numpy(), want_dB.astype(np.float32))
    assert np.array_equal(dv.cpu().numpy(), want_dv.astype(np.float32))


@pytest.mark.parametrize("N,b_dtype,knobs", [
    (64, torch.float32, None),
    (64, torch.bfloat16, None),                # short rows + bf16 B: the row-batch form
    (100, torch.float32, {"thread_nz": 1}),    # carries at nearly every chunk, a partial lane
    (32, torch.bfloat16, {"block_numer": 1, "block_denom": 1000000}),
])
def test_fused_backward_rowmap_equals_gathered(gpu, N, b_dtype, knobs):
    """dC through a row map (A's row r reads dP[map[r]]) gives what the plain entry point
    gives on the gathered copy dP[map]: dvals bit for bit (same products in the same
    order), dB up to the order of its atomic carries; both exact on integer operands."""
    m = _short_row_matrix(N + 3)
    rng = np.random.default_rng(N + 1)
    S = 37
    gmap = np.sort(rng.integers(0, S, size=m.nrow)).astype(np.int32)  # e.g. a pooling index
    dP = rng.integers(-3, 4, size=(S, N)).astype(np.float32)
    B = rng.integers(-3, 4, size=(m.ncol, N)).astype(np.float32)
    A = m.to_device(gpu)
    T = ops.csr_transpose(A)
    Bg = torch.from_numpy(B).to(gpu).to(b_dtype)
    dPg = torch.from_numpy(dP).to(gpu)
    mapg = torch.from_numpy(gmap).to(gpu)
    dB1, dv1 = ops.spmm_backward(T, A.vals, dPg, Bg, knobs=knobs, dc_map=mapg)
    dB2, dv2 = ops.spmm_backward(T, A.vals, dPg[mapg.long()].contiguous(), Bg, knobs=knobs)
    assert torch.equal(dv1, dv2)
    assert torch.equal(dB1, dB2)  # integer operands: every sum is exact
    rows = m.row_ids()
    want = np.zeros((m.ncol, N))
    np.add.at(want, m.colind, m.vals[:, None].astype(np.float64) * dP[gmap[rows]])
    assert np.array_equal(dB1.cpu().numpy(), want.astype(np.float32))
    with pytest.raises(ValueError):
        ops.spmm_backward(T, A.vals, dPg, Bg, dc_map=mapg[:-1])
