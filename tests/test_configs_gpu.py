"""Parity at the full sizes of the five measured workloads, through the C ABI,
against the CPU oracle — the check mode of the reference's harnesses
(test-cg/src/test.cu:52-82 checks every SpMM variant against
spmm_reference_host on the whole matrix; new-cg/src/main.cu:33-58 checks
segment_coo against segment_coo_sequencial), run here on the workloads
themselves rather than on shrunk stand-ins:

  Poisson     16384 x 16384, nnz/row ~ Poisson(16), N = 16
  power-law   262144 x 262144 Zipf(2.1), 4,194,295 nnz, N in {4, 16, 64, 128}, all
              variants plus AUTO and the committed tuned points (fp32 and bf16 B);
              segment_coo over the same degrees (E = 4,194,295, N = 64, fp32 / bf16 src)
  banded      120000 x 120000 of 3 x 3 blocks, N in {64, 128}, B bf16 / fp16 / fp32
  GCN         D^-1/2 (A + I) D^-1/2, 131072 nodes, N = 128, X fp32 / bf16: forward,
              fused backward (dX = A^T dC, dvals = SDDMM(dC, X)), mean-pool backward
  batched     256 graphs (2,394,488 nodes, 9,580,737 nnz), bf16 features, N = 64:
              forward, pooling and backward, eager and hipGraph replays

Two checks per case:
  * integer operands (values and dense entries in {-2..2}; every partial sum is an
    integer below 2^24, so every summation order is exact): the GPU result equals
    the fp32 oracle bit for bit — row and segment boundaries, carries, empty rows;
  * U(-1, 1) operands: normalized error max |C - C64| / (|A| |B|) <= 1e-6 against
    the float64 oracle on the operand as stored (bf16 / fp16 rounded), and
    <= 8e-3 against the unrounded fp32 operand for 16-bit storage.
"""
import numpy as np
import pytest
import torch

from segscan_amd import autotune, batch, ops, synth

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-6
HALF_TOL = 8e-3

ALL_SPMM = ["plain_ebpr", "plain_rbpr", "plain_ebsr", "plain_rbsr", "ebpr", "rbpr", "ebsr", "rbsr", "auto"]
SEG_VARIANTS = ["segment_coo", "segment_coo_new", "segment_coo_sr", "segment_coo_noshmem_sr",
                "segment_coo_noshmem_lessatom_sr", "segment_coo_sorted_sr", "auto"]


def _int_copy(m: synth.HostCsr, seed: int) -> synth.HostCsr:
    rng = np.random.default_rng(seed)
    vals = rng.integers(-2, 3, size=m.nnz, dtype=np.int8).astype(np.float32)
    return synth.HostCsr(m.nrow, m.ncol, m.rowptr, m.colind, vals)


def _int_dense(rng, rows, n):
    return rng.integers(-2, 3, size=(rows, n), dtype=np.int8).astype(np.float32)


def _run_spmm(A, B, variant, knobs=None):
    C = torch.full((A.nrow, B.size(1)), float("nan"), dtype=torch.float32, device=B.device)
    ops.spmm(A, B, variant=variant, knobs=knobs, out=C)
    return C.cpu().numpy()


def _tuned_point(stats, ncol, dtype, N):
    """The committed per-matrix tuning result that bench.py runs, if any."""
    hit = autotune.lookup(autotune.load_table(), "spmm", dtype, N, stats, ncol)
    if hit is None or "|m" not in hit[2]:
        return None
    return hit[0], hit[1]


@pytest.fixture(scope="module")
def poisson():
    return synth.poisson_csr(seed=0)


@pytest.fixture(scope="module")
def powerlaw():
    return synth.powerlaw_csr(seed=0)


@pytest.fixture(scope="module")
def banded():
    return synth.banded_block_csr(seed=0)


# ---- Poisson 16384^2, N = 16 ----------------------------------------------------------------

def test_poisson_all_variants(gpu, oracle, poisson):
    m, N = poisson, 16
    print(f"poisson {m.nrow}x{m.ncol} nnz={m.nnz} N={N}", flush=True)
    rng = np.random.default_rng(1)
    mi = _int_copy(m, 2)
    Bi = _int_dense(rng, m.ncol, N)
    Ci_ref = oracle.spmm_f32(mi.rowptr, mi.colind, mi.vals, Bi)
    Bf = rng.uniform(-1, 1, size=(m.ncol, N)).astype(np.float32)
    C64, D = oracle.spmm_f64(m.rowptr, m.colind, m.vals, Bf)
    Ai, Af = mi.to_device(gpu), m.to_device(gpu)
    Bi_d, Bf_d = torch.from_numpy(Bi).to(gpu), torch.from_numpy(Bf).to(gpu)
    for v in ALL_SPMM:
        assert np.array_equal(_run_spmm(Ai, Bi_d, v), Ci_ref), v
        err = oracle.normalized_error(_run_spmm(Af, Bf_d, v), C64, D)
        assert err <= FP32_TOL, (v, err)


def test_poisson_reference_check_mode(gpu, oracle, poisson):
    """The reference harness's own inputs: values all 1 (vals NULL) and B all 1, checked
    with its check_result rule (1e-2 relative, test-cg/include/util/check.cuh:116-136);
    the exact answer is C[i][j] = deg(i)."""
    m, N = poisson, 16
    A = synth.HostCsr(m.nrow, m.ncol, m.rowptr, m.colind, None).to_device(gpu)
    B = torch.ones((m.ncol, N), dtype=torch.float32, device=gpu)
    deg = np.repeat(m.degrees().astype(np.float32)[:, None], N, axis=1)
    for v in ALL_SPMM:
        C = _run_spmm(A, B, v)
        assert oracle.check_result(C, deg) == 0, v
        assert np.array_equal(C, deg), v


# ---- power-law 262144^2, N in {4, 16, 64, 128} ------------------------------------------------

@pytest.mark.parametrize("N", [4, 16, 64, 128])
def test_powerlaw_all_variants(gpu, oracle, powerlaw, N):
    m = powerlaw
    print(f"powerlaw {m.nrow}x{m.ncol} nnz={m.nnz} N={N}", flush=True)
    stats = autotune.row_stats(m.rowptr)
    rng = np.random.default_rng(100 + N)
    mi = _int_copy(m, 3)
    Bi = _int_dense(rng, m.ncol, N)
    Ci_ref = oracle.spmm_f32(mi.rowptr, mi.colind, mi.vals, Bi)
    Bf = rng.uniform(-1, 1, size=(m.ncol, N)).astype(np.float32)
    C64, D = oracle.spmm_f64(m.rowptr, m.colind, m.vals, Bf)
    Ai, Af = mi.to_device(gpu), m.to_device(gpu)
    Bi_d, Bf_d = torch.from_numpy(Bi).to(gpu), torch.from_numpy(Bf).to(gpu)
    cases = [(v, None) for v in ALL_SPMM]
    tp = _tuned_point(stats, m.ncol, torch.float32, N)
    if tp is not None:
        cases.append(tp)
    for v, k in cases:
        assert np.array_equal(_run_spmm(Ai, Bi_d, v, k), Ci_ref), (v, k)
        err = oracle.normalized_error(_run_spmm(Af, Bf_d, v, k), C64, D)
        assert err <= FP32_TOL, (v, k, err)


def test_powerlaw_bf16_tuned_point(gpu, oracle, powerlaw):
    """bf16 B at N = 64: AUTO and the committed tuned point (the bench's bf16 leg)."""
    m, N = powerlaw, 64
    stats = autotune.row_stats(m.rowptr)
    rng = np.random.default_rng(7)
    cases = [("auto", None), ("ebsr", None)]
    tp = _tuned_point(stats, m.ncol, torch.bfloat16, N)
    if tp is not None:
        cases.append(tp)
    Af = m.to_device(gpu)
    Bf = rng.uniform(-1, 1, size=(m.ncol, N)).astype(np.float32)
    Bq_d = torch.from_numpy(Bf).to(gpu).to(torch.bfloat16)
    Bq = Bq_d.float().cpu().numpy()
    C64q, Dq = oracle.spmm_f64(m.rowptr, m.colind, m.vals, Bq)
    C64, D = oracle.spmm_f64(m.rowptr, m.colind, m.vals, Bf)
    mi = _int_copy(m, 4)
    Bi = _int_dense(rng, m.ncol, N)
    Ci_ref = oracle.spmm_f32(mi.rowptr, mi.colind, mi.vals, Bi)
    Ai, Bi_d = mi.to_device(gpu), torch.from_numpy(Bi).to(gpu).to(torch.bfloat16)
    for v, k in cases:
        C = _run_spmm(Af, Bq_d, v, k)
        assert oracle.normalized_error(C, C64q, Dq) <= FP32_TOL, (v, k)
        assert oracle.normalized_error(C, C64, D) <= HALF_TOL, (v, k)
        assert np.array_equal(_run_spmm(Ai, Bi_d, v, k), Ci_ref), (v, k)


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_powerlaw_segment_coo(gpu, oracle, powerlaw, dtype):
    """segment_coo over the power-law degrees: E = 4,194,295 sorted indices, src E x 64."""
    m, N = powerlaw, 64
    idx = synth.segment_index_from_degrees(m.degrees())
    E = idx.size
    print(f"segment_coo E={E} N={N} {dtype}", flush=True)
    rng = np.random.default_rng(9)
    idx_d = torch.from_numpy(idx).to(gpu)
    src_i = _int_dense(rng, E, N)
    ref_i = oracle.segment_coo_f32(src_i, idx, m.nrow)
    src_i_d = torch.from_numpy(src_i).to(gpu).to(dtype)
    del src_i
    src_f = rng.uniform(-1, 1, size=(E, N)).astype(np.float32)
    src_f_d = torch.from_numpy(src_f).to(gpu).to(dtype)
    src_q = src_f_d.float().cpu().numpy()
    ref64, absd = oracle.segment_coo_f64(src_q, idx, m.nrow)
    del src_q
    if dtype != torch.float32:
        ref64_u, absd_u = oracle.segment_coo_f64(src_f, idx, m.nrow)
    del src_f
    out = torch.empty((m.nrow, N), dtype=torch.float32, device=gpu)
    for v in SEG_VARIANTS:
        out.fill_(float("nan"))
        ops.segment_coo(src_i_d, idx_d, m.nrow, variant=v, out=out)
        assert np.array_equal(out.cpu().numpy(), ref_i), v
        out.fill_(float("nan"))
        ops.segment_coo(src_f_d, idx_d, m.nrow, variant=v, out=out)
        o = out.cpu().numpy()
        assert oracle.normalized_error(o, ref64, absd) <= FP32_TOL, v
        if dtype != torch.float32:
            assert oracle.normalized_error(o, ref64_u, absd_u) <= HALF_TOL, v


# ---- banded 120000^2, 16-bit B ------------------------------------------------------------

@pytest.mark.parametrize("N", [64, 128])
def test_banded_mixed_precision(gpu, oracle, banded, N):
    m = banded
    print(f"banded {m.nrow}x{m.ncol} nnz={m.nnz} N={N}", flush=True)
    rng = np.random.default_rng(20 + N)
    A = m.to_device(gpu)
    Bf = rng.uniform(-1, 1, size=(m.ncol, N)).astype(np.float32)
    C64, D = oracle.spmm_f64(m.rowptr, m.colind, m.vals, Bf)
    mi = _int_copy(m, 5)
    Ai = mi.to_device(gpu)
    Bi = _int_dense(rng, m.ncol, N)
    Ci_ref = oracle.spmm_f32(mi.rowptr, mi.colind, mi.vals, Bi)
    for dtype in (torch.bfloat16, torch.float16, torch.float32):
        Bq_d = torch.from_numpy(Bf).to(gpu).to(dtype)
        Bq = Bq_d.float().cpu().numpy()
        C64q, Dq = (C64, D) if dtype == torch.float32 else oracle.spmm_f64(m.rowptr, m.colind, m.vals, Bq)
        Bi_d = torch.from_numpy(Bi).to(gpu).to(dtype)
        for v in ("auto", "rbsr", "plain_rbsr", "ebsr", "rbpr"):
            C = _run_spmm(A, Bq_d, v)
            assert oracle.normalized_error(C, C64q, Dq) <= FP32_TOL, (dtype, v)
            assert oracle.normalized_error(C, C64, D) <= HALF_TOL, (dtype, v)
            assert np.array_equal(_run_spmm(Ai, Bi_d, v), Ci_ref), (dtype, v)


# ---- GCN layer, 131072 nodes, N = 128 --------------------------------------------------------

@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_gcn_forward_backward(gpu, oracle, dtype):
    m, N = synth.gcn_normalized(seed=0), 128
    print(f"gcn {m.nrow} nodes nnz={m.nnz} N={N} {dtype}", flush=True)
    rng = np.random.default_rng(30)
    A = m.to_device(gpu)
    X_d = torch.from_numpy(rng.standard_normal((m.ncol, N)).astype(np.float32)).to(gpu).to(dtype)
    Xq = X_d.float().cpu().numpy()
    dCh = rng.standard_normal((m.nrow, N)).astype(np.float32)
    dC = torch.from_numpy(dCh).to(gpu)
    # forward
    C64, D = oracle.spmm_f64(m.rowptr, m.colind, m.vals, Xq)
    assert oracle.normalized_error(_run_spmm(A, X_d, "auto"), C64, D) <= FP32_TOL
    # fused backward: dX = A^T dC, dvals = <dC[row], X[col]>
    colptr, rowind, perm = oracle.transpose(m.nrow, m.ncol, m.rowptr, m.colind)
    dX64, dXabs = oracle.spmm_f64(colptr, rowind, m.vals[perm], dCh)
    dv64, dvabs = oracle.sddmm_f64(m.rowptr, m.colind, dCh, Xq)
    T = ops.csr_transpose(A)
    dX, dvals = ops.spmm_backward(T, A.vals, dC, X_d)
    assert oracle.normalized_error(dX.cpu().numpy(), dX64, dXabs) <= FP32_TOL
    assert oracle.normalized_error(dvals.cpu().numpy(), dv64, dvabs) <= FP32_TOL
    # values in the transpose's order (a fixed adjacency)
    dX2, _ = ops.spmm_backward(T, None, dC, X_d, want_dvals=False, vals_t=T.t.vals)
    assert oracle.normalized_error(dX2.cpu().numpy(), dX64, dXabs) <= FP32_TOL
    # mean-pool backward over 512-node groups: dH[e] = dP[g(e)] / |g|
    gid = (np.arange(m.nrow) // 512).astype(np.int32)
    ng = int(gid[-1]) + 1
    dPh = rng.standard_normal((ng, N)).astype(np.float32)
    counts = ops.segment_counts(torch.from_numpy(gid).to(gpu), ng)
    dH = ops.segment_gather(torch.from_numpy(dPh).to(gpu), torch.from_numpy(gid).to(gpu), counts).cpu().numpy()
    cnt = np.bincount(gid, minlength=ng).astype(np.float32)
    assert np.array_equal(counts.cpu().numpy(), cnt.astype(np.int32))
    assert np.allclose(dH, dPh[gid] / cnt[gid][:, None], rtol=5e-7, atol=0)


# ---- batched: 256 uneven graphs, bf16 features, N = 64 ---------------------------------------

def _batched_reference(oracle, csr, offsets, Xq, dP):
    gi = np.repeat(np.arange(offsets.size - 1), np.diff(offsets)).astype(np.int32)
    H64, HD = oracle.spmm_f64(csr.rowptr, csr.colind, csr.vals, Xq)
    P64 = np.zeros((offsets.size - 1, Xq.shape[1]))
    PD = np.zeros_like(P64)
    np.add.at(P64, gi, H64)
    np.add.at(PD, gi, HD)
    dHt = dP[gi]
    colptr, rowind, perm = oracle.transpose(csr.nrow, csr.ncol, csr.rowptr, csr.colind)
    dX64, dXabs = oracle.spmm_f64(colptr, rowind, csr.vals[perm], dHt)
    dv64, dvabs = oracle.sddmm_f64(csr.rowptr, csr.colind, dHt, Xq)
    return (H64, HD), (P64, PD), (dX64, dXabs), (dv64, dvabs)


def _batched_check(oracle, step, ref):
    for name, got, (r, d) in zip(("H", "P", "dX", "dvals"), (step.H, step.P, step.dX, step.dvals), ref):
        err = oracle.normalized_error(got.cpu().numpy(), r, d)
        assert err <= FP32_TOL, (name, err)


@pytest.mark.parametrize("single_launch", [True, False])
def test_batched_256_graphs_bf16(gpu, oracle, single_launch):
    """The default step (one launch of the batch plan) and the kernel-by-kernel form."""
    csr, offsets = synth.batched_graphs(seed=0)
    print(f"batched {offsets.size - 1} graphs nodes={csr.nrow} nnz={csr.nnz} single_launch={single_launch}",
          flush=True)
    b = batch.make_batch(csr, offsets, gpu)
    N = 64
    rng = np.random.default_rng(40)
    X = torch.empty((csr.nrow, N), dtype=torch.bfloat16, device=gpu)
    dP = torch.empty((b.n_graphs, N), dtype=torch.float32, device=gpu)

    def new_inputs():
        X.copy_(torch.from_numpy(rng.uniform(-1, 1, size=(csr.nrow, N)).astype(np.float32)).to(gpu))
        dP.copy_(torch.from_numpy(rng.standard_normal((b.n_graphs, N)).astype(np.float32)).to(gpu))
        return X.float().cpu().numpy(), dP.cpu().numpy()

    step = batch.BatchedGcnStep(b, N, dtype=torch.bfloat16, single_launch=single_launch)
    Xq, dPh = new_inputs()
    step.run(X, dP)
    torch.cuda.synchronize()
    _batched_check(oracle, step, _batched_reference(oracle, csr, offsets, Xq, dPh))
    g = step.capture(X, dP)
    Xq, dPh = new_inputs()
    ref = _batched_reference(oracle, csr, offsets, Xq, dPh)
    for rep in range(3):
        # eager work between replays: outputs scribbled over, then one replay
        step.H.fill_(123.0)
        step.P.fill_(5.0)
        step.dX.fill_(-7.0)
        step.dvals.fill_(3.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        print(f"replay {rep}", flush=True)
        _batched_check(oracle, step, ref)
