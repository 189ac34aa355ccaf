"""SpMM parity on the GPU: every variant (plain and tunable) and a spread of
knobs against the CPU oracle (the reference's host check path restated).

* integer-valued inputs (values and B in {-2..2}): every summation order is
  exact, so the GPU result must equal the oracle bit for bit — this pins row
  boundaries, column indexing, empty rows and the carry / atomic write paths;
* U(-1, 1) inputs: normalized error max|C - C64| / (|A||B|) <= 1e-6 against the
  float64 oracle (fp32 storage), <= 1e-6 against the float64 oracle on the
  rounded operand and <= 8e-3 against the unrounded one for bf16 / fp16 B.
"""
import numpy as np
import pytest
import torch

from segscan_amd import _abi, ops, synth

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-6
HALF_TOL = 8e-3

CONFIGS = [
    ("plain_ebpr", None),
    ("plain_rbpr", None),
    ("plain_ebsr", None),
    ("plain_rbsr", None),
    ("ebpr", None),
    ("rbpr", None),
    ("ebsr", None),
    ("rbsr", None),
    ("auto", None),
    ("ebpr", dict(group_factor=6, tile_factor=6, thread_per_block=256)),
    ("ebpr", dict(group_factor=2, tile_factor=4, block_numer=1, block_denom=64)),
    ("ebpr", dict(group_factor=0, tile_factor=2)),
    ("ebpr", dict(group_factor=4, tile_factor=5, thread_per_block=512)),
    ("rbpr", dict(group_factor=2, tile_factor=5)),
    ("rbpr", dict(group_factor=6, tile_factor=6)),
    ("rbpr", dict(group_factor=3, tile_factor=3, block_numer=1, block_denom=8)),
    ("ebsr", dict(group_factor=4, coarsen_factor=4, thread_nz=16, write_policy=2)),
    ("ebsr", dict(group_factor=6, coarsen_factor=1, thread_nz=4, write_policy=1)),
    ("ebsr", dict(group_factor=3, coarsen_factor=8, thread_nz=8, write_policy=3)),
    ("ebsr", dict(group_factor=0, coarsen_factor=2, thread_nz=7, write_policy=2)),
    ("ebsr", dict(group_factor=5, coarsen_factor=2, thread_nz=3, write_policy=2, block_numer=1, block_denom=1000)),
    ("rbsr", dict(coarsen_factor=2, tile_factor=3, thread_per_block=128)),
    ("rbsr", dict(coarsen_factor=4, tile_factor=6, thread_per_block=64, block_numer=1, block_denom=7)),
]

NS = [1, 3, 4, 16, 33, 64]


def _ids(c):
    v, k = c
    return v + ("" if not k else "-" + "-".join(f"{a[:2]}{b}" for a, b in k.items()))


@pytest.fixture(scope="module")
def mats():
    """Small matrices with empty rows, a dense row and a power-law tail."""
    m1 = synth.small_random_csr(300, 200, 0.05, seed=1, values="int", empty_rows=0.2)
    m2 = synth.powerlaw_csr(M=2000, K=1500, alpha=2.1, max_deg=1500, target_nnz=20000, seed=2, values="int")
    # a matrix with long runs of empty rows and a single dense row
    deg = np.zeros(500, dtype=np.int64)
    deg[17] = 400
    deg[250:260] = 3
    deg[499] = 5
    rng = np.random.default_rng(3)
    m3 = synth.csr_from_degrees(rng, deg, 400, values="int")
    return [m1, m2, m3]


def _int_dense(rng, rows, n):
    return rng.integers(-2, 3, size=(rows, n)).astype(np.float32)


@pytest.mark.parametrize("cfg", CONFIGS, ids=_ids)
def test_spmm_integer_exact(gpu, oracle, mats, cfg):
    variant, knobs = cfg
    rng = np.random.default_rng(11)
    for mi, m in enumerate(mats):
        A = m.to_device(gpu)
        for N in NS:
            B = _int_dense(rng, m.ncol, N)
            ref = oracle.spmm_f32(m.rowptr, m.colind, m.vals, B)
            C = ops.spmm(A, torch.from_numpy(B).to(gpu), variant=variant, knobs=knobs)
            got = C.cpu().numpy()
            assert np.array_equal(got, ref), f"matrix {mi} N={N}: max diff {np.abs(got - ref).max()}"


@pytest.mark.parametrize("cfg", CONFIGS, ids=_ids)
def test_spmm_float_normalized(gpu, oracle, cfg):
    variant, knobs = cfg
    m = synth.powerlaw_csr(M=3000, K=2500, alpha=2.1, max_deg=2500, target_nnz=40000, seed=5)
    A = m.to_device(gpu)
    rng = np.random.default_rng(5)
    for N in (4, 64):
        B = rng.uniform(-1, 1, size=(m.ncol, N)).astype(np.float32)
        C64, D = oracle.spmm_f64(m.rowptr, m.colind, m.vals, B)
        C = ops.spmm(A, torch.from_numpy(B).to(gpu), variant=variant, knobs=knobs).cpu().numpy()
        err = oracle.normalized_error(C, C64, D)
        assert err <= FP32_TOL, f"N={N} normalized error {err}"


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16])
@pytest.mark.parametrize("variant", ["plain_ebpr", "ebpr", "rbpr", "ebsr", "rbsr", "auto"])
def test_spmm_half_storage(gpu, oracle, dtype, variant):
    m = synth.powerlaw_csr(M=3000, K=2500, alpha=2.1, max_deg=2500, target_nnz=40000, seed=6)
    A = m.to_device(gpu)
    rng = np.random.default_rng(6)
    for N in (8, 64, 128):
        B32 = rng.uniform(-1, 1, size=(m.ncol, N)).astype(np.float32)
        Bh = torch.from_numpy(B32).to(dtype)
        B_rounded = Bh.float().numpy()
        C = ops.spmm(A, Bh.to(gpu), variant=variant).cpu().numpy()
        C64r, Dr = oracle.spmm_f64(m.rowptr, m.colind, m.vals, B_rounded)
        assert oracle.normalized_error(C, C64r, Dr) <= FP32_TOL
        C64, D = oracle.spmm_f64(m.rowptr, m.colind, m.vals, B32)
        assert oracle.normalized_error(C, C64, D) <= HALF_TOL


@pytest.mark.parametrize("variant", ["plain_ebpr", "rbpr", "ebsr", "rbsr", "auto"])
def test_spmm_accumulate_and_strides(gpu, oracle, mats, variant):
    rng = np.random.default_rng(21)
    m = mats[1]
    A = m.to_device(gpu)
    N = 20
    # B and C as column slices of wider row-major buffers (ld > N)
    Bw = torch.from_numpy(_int_dense(rng, m.ncol, 32)).to(gpu)
    Cw = torch.from_numpy(_int_dense(rng, m.nrow, 40)).to(gpu)
    C0 = Cw[:, 4:4 + N].cpu().numpy().copy()
    B = Bw[:, 8:8 + N]
    ops.spmm(A, B, variant=variant, out=Cw[:, 4:4 + N], accumulate=True)
    ref = oracle.spmm_f32(m.rowptr, m.colind, m.vals, B.cpu().numpy(), C=C0.copy())
    assert np.array_equal(Cw[:, 4:4 + N].cpu().numpy(), ref)
    # columns outside the slice are untouched
    assert np.array_equal(Cw[:, :4].cpu().numpy(), Cw[:, :4].cpu().numpy())


def test_spmm_all_ones_known_answer(gpu):
    """The reference harness fills values and B with ones (dataloader.hpp:154-155,
    test.cu:42), so C[i][j] = deg(i)."""
    m = synth.powerlaw_csr(M=5000, K=5000, alpha=2.1, max_deg=3000, target_nnz=60000, seed=9)
    A = m.to_device(gpu)
    A.vals = None  # NULL values read as 1
    B = torch.ones((m.ncol, 64), device=gpu)
    deg = m.degrees().astype(np.float32)
    for variant in ("plain_ebpr", "plain_rbpr", "plain_ebsr", "plain_rbsr", "ebpr", "rbpr", "ebsr", "rbsr", "auto"):
        C = ops.spmm(A, B, variant=variant).cpu().numpy()
        assert np.array_equal(C, np.repeat(deg[:, None], 64, axis=1)), variant


def test_spmm_empty_and_degenerate(gpu):
    dev = gpu
    # nnz = 0: C must be zero-filled
    A = ops.CsrMatrix.from_numpy(7, 5, np.zeros(8, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32), device=dev)
    B = torch.randn(5, 9, device=dev)
    for variant in ("plain_ebpr", "ebsr", "rbsr", "rbpr", "auto"):
        C = torch.full((7, 9), 3.0, device=dev)
        ops.spmm(A, B, variant=variant, out=C)
        assert torch.count_nonzero(C).item() == 0, variant
    # M = 0 and N = 0 are no-ops
    A0 = ops.CsrMatrix.from_numpy(0, 5, np.zeros(1, np.int32), np.zeros(0, np.int32), None, device=dev)
    ops.spmm(A0, B)
    A1 = ops.CsrMatrix.from_numpy(2, 5, np.array([0, 1, 2], np.int32), np.array([0, 4], np.int32), None, device=dev)
    ops.spmm(A1, torch.zeros(5, 0, device=dev))


def test_check_csr_flags(gpu):
    m = synth.small_random_csr(50, 40, 0.2, seed=4)
    assert ops.check_csr(m.to_device(gpu)) == 0
    bad = synth.small_random_csr(50, 40, 0.2, seed=4)
    bad.colind = bad.colind.copy()
    bad.colind[3] = 40  # out of range
    assert ops.check_csr(bad.to_device(gpu)) & 2
