"""Long runs of empty rows and segments, against the oracle. Under the owner write
policy a kernel zeroes only the empty rows that share a 16-row block with a row that
has entries; whole empty blocks are zeroed by the fix-up pass with its whole grid, so
no run, however long, falls on one group. Cases: a CSR with 20k leading, 400k interior
and 600k trailing empty rows; a segment_coo over 4,194,304 segments whose index uses
only the first 1 %, and one whose few used segments lie a million apart; the fused
backward over a matrix whose used columns (dB rows) sit in two narrow bands of 1.07M.
Operands are small integers, so every variant must match the oracle bit for bit.
Accumulating calls must leave empty rows as they were."""
import time

import numpy as np
import pytest
import torch

from segscan_amd import ops, synth

pytestmark = pytest.mark.gpu

OWNER = 2
SPMM_CASES = ["auto", "ebsr", "plain_ebsr", "ebpr", "rbsr", "rbpr",
              ("ebsr", dict(group_factor=5, coarsen_factor=2, thread_nz=8, write_policy=OWNER)),
              ("ebsr", dict(group_factor=4, coarsen_factor=4, thread_nz=4, write_policy=OWNER)),
              ("ebsr", dict(group_factor=2, coarsen_factor=1, thread_nz=64, write_policy=OWNER))]
SEG_CASES = ["auto", "segment_coo", "segment_coo_new", "segment_coo_sr", "segment_coo_noshmem_sr",
             "segment_coo_noshmem_lessatom_sr", "segment_coo_sorted_sr",
             ("segment_coo_sr", dict(write_policy=OWNER)),
             ("segment_coo_sr", dict(write_policy=OWNER, group_factor=4, thread_nz=8))]


def _case(c):
    return c if isinstance(c, tuple) else (c, None)


def _runs_csr(seed):
    """1,070,000 rows: 20k empty, 25k Zipf rows, 400k empty, 25k Zipf rows, 600k empty.
    The Zipf rows hold short interior gaps too."""
    rng = np.random.default_rng(seed)
    M, K = 1_070_000, 30_000
    deg = np.zeros(M, np.int64)
    for start in (20_000, 445_000):
        body = np.clip(rng.zipf(2.0, size=25_000), 1, 3000)
        body[rng.random(25_000) < 0.1] = 0
        body[1000:1300] = 0
        deg[start:start + 25_000] = body
    return synth.csr_from_degrees(rng, deg, K, values="int")


def _time(fn):
    torch.cuda.synchronize()
    t = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t) * 1e3


@pytest.fixture(scope="module")
def runs_csr():
    return _runs_csr(1)


def test_spmm_long_empty_runs(gpu, oracle, runs_csr):
    m = runs_csr
    print(f"csr {m.nrow}x{m.ncol} nnz={m.nnz}", flush=True)
    A = m.to_device(gpu)
    rng = np.random.default_rng(2)
    N = 64
    Bh = rng.integers(-2, 3, size=(m.ncol, N)).astype(np.float32)
    B = torch.from_numpy(Bh).to(gpu)
    ref = oracle.spmm_f32(m.rowptr, m.colind, m.vals, Bh)
    C = torch.empty((m.nrow, N), device=gpu)
    for case in SPMM_CASES:
        v, k = _case(case)
        C.fill_(float("nan"))
        ms = _time(lambda: ops.spmm(A, B, variant=v, knobs=k, out=C))
        print(f"  {v} {k}: {ms:.2f} ms", flush=True)
        assert np.array_equal(C.cpu().numpy(), ref), (v, k)


def test_spmm_accumulate_keeps_empty_rows(gpu, oracle, runs_csr):
    m = runs_csr
    A = m.to_device(gpu)
    rng = np.random.default_rng(3)
    N = 16
    Bh = rng.integers(-2, 3, size=(m.ncol, N)).astype(np.float32)
    B = torch.from_numpy(Bh).to(gpu)
    C0 = rng.integers(-3, 4, size=(m.nrow, N)).astype(np.float32)
    ref = oracle.spmm_f32(m.rowptr, m.colind, m.vals, Bh, C=C0.copy())
    for k in (dict(write_policy=OWNER), dict(write_policy=OWNER, group_factor=2, thread_nz=32)):
        C = torch.from_numpy(C0).to(gpu)
        ops.spmm(A, B, variant="ebsr", knobs=k, out=C, accumulate=True)
        assert np.array_equal(C.cpu().numpy(), ref), k


def test_backward_empty_column_bands(gpu, oracle, runs_csr):
    """dB = A^T dC where A's columns map to 10k..25k and 515k..530k of 1.07M: the
    transpose has 10k leading, ~490k interior and ~540k trailing empty rows."""
    m = runs_csr
    col = m.colind.astype(np.int64)
    col = (col + 10_000 + (col >= 15_000) * 490_000).astype(np.int32)
    A = synth.HostCsr(m.nrow, m.nrow, m.rowptr, col, m.vals).to_device(gpu)
    rng = np.random.default_rng(4)
    N = 64
    dCh = rng.integers(-2, 3, size=(m.nrow, N)).astype(np.float32)
    Bh = rng.integers(-2, 3, size=(m.nrow, N)).astype(np.float32)
    T = ops.csr_transpose(A)
    dC, B = torch.from_numpy(dCh).to(gpu), torch.from_numpy(Bh).to(gpu)
    colptr, rowind, perm = oracle.transpose(m.nrow, m.nrow, m.rowptr, col)
    ref = oracle.spmm_f32(colptr, rowind, m.vals[perm], dCh)
    dv64, _ = oracle.sddmm_f64(m.rowptr, col, dCh, Bh)
    dB = torch.full((m.nrow, N), float("nan"), device=gpu)
    dB, dv = ops.spmm_backward(T, A.vals, dC, B, dB=dB)
    assert np.array_equal(dB.cpu().numpy(), ref)
    assert np.array_equal(dv.cpu().numpy(), dv64.astype(np.float32))
    # accumulating: empty dB rows keep their values
    base = rng.integers(-3, 4, size=(m.nrow, N)).astype(np.float32)
    dB = torch.from_numpy(base).to(gpu)
    ops.spmm_backward(T, A.vals, dC, B, want_dvals=False, dB=dB, accumulate=True)
    assert np.array_equal(dB.cpu().numpy(), base + ref)


def _seg_check(gpu, oracle, idx, dst_len, dtype, seed):
    rng = np.random.default_rng(seed)
    E, N = idx.size, 64
    print(f"segment_coo E={E} dst_len={dst_len} {dtype}", flush=True)
    src = rng.integers(-2, 3, size=(E, N)).astype(np.float32)
    ref = oracle.segment_coo_f32(src, idx, dst_len)
    src_d = torch.from_numpy(src).to(gpu).to(dtype)
    idx_d = torch.from_numpy(idx).to(gpu)
    out = torch.empty((dst_len, N), device=gpu)
    for case in SEG_CASES:
        v, k = _case(case)
        out.fill_(float("nan"))
        ms = _time(lambda: ops.segment_coo(src_d, idx_d, dst_len, variant=v, knobs=k, out=out))
        print(f"  {v} {k}: {ms:.2f} ms", flush=True)
        assert np.array_equal(out.cpu().numpy(), ref), (v, k)
    # accumulating owner call: untouched segments keep their values
    base = rng.integers(-3, 4, size=(dst_len, 8)).astype(np.float32)
    out8 = torch.from_numpy(base).to(gpu)
    ops.segment_coo(src_d[:, :8].contiguous(), idx_d, dst_len, variant="segment_coo_sr",
                    knobs=dict(write_policy=OWNER), out=out8, accumulate=True)
    assert np.array_equal(out8.cpu().numpy(), base + ref[:, :8])


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_segment_coo_first_percent_of_4m(gpu, oracle, dtype):
    """dst_len = 4,194,304; the sorted index uses segments 1000..42943 (1 %)."""
    rng = np.random.default_rng(5)
    dst_len = 4_194_304
    used = np.arange(1000, 1000 + dst_len // 100)
    idx = np.repeat(used, rng.integers(0, 50, size=used.size)).astype(np.int32)
    _seg_check(gpu, oracle, idx, dst_len, dtype, 6)


def test_segment_coo_few_segments_far_apart(gpu, oracle):
    """Five used segments of 4,194,304, a million apart, 20k entries each."""
    dst_len = 4_194_304
    idx = np.repeat(np.array([3, 1_000_017, 2_000_000, 3_000_031, 4_194_303]), 20_000).astype(np.int32)
    _seg_check(gpu, oracle, idx, dst_len, torch.float32, 7)
