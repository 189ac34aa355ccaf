"""Seeded knob-space sweep on the GPU: SpMM (every variant), segment_coo (the six
reference launchers and AUTO) and the fused backward, each with knobs, shapes and
storage types drawn from a fixed seed, checked bit for bit against the CPU oracle.

Every operand is a small integer (A's values and the dense rows in {-3..3}), so each
partial sum is exact in fp32 in any order and in bf16 / fp16 storage; the GPU result
must then equal the oracle exactly, whatever the lane layout, chunking or write
policy. A combination the library does not compile is rejected with
SEGSCAN_EUNSUPPORTED (or EINVAL for an inconsistent knob set) before anything is
launched; that is allowed, anything else is a failure. The cases are fixed by the
seed, one test item each, so a failure names its case.
"""
import numpy as np
import pytest
import torch

from segscan_amd import _abi, ops, synth

pytestmark = pytest.mark.gpu

SEED = 20240611
NS = [1, 2, 3, 4, 7, 8, 16, 31, 33, 64, 65, 100, 128, 200]
DTYPES = {"f32": torch.float32, "bf16": torch.bfloat16, "f16": torch.float16}
TPB = [64, 128, 256, 512, 1024]
DENOMS = [1, 8, 64, 1000, 100000]


def _int_values(rng, n):
    return rng.integers(-3, 4, size=n).astype(np.float32)


def _matrix(rng, kind):
    """Small CSR with integer values: random (with empty rows), power-law rows,
    one dense row among empty ones, or no entries at all."""
    if kind == "random":
        M, K = int(rng.integers(1, 700)), int(rng.integers(1, 600))
        m = synth.small_random_csr(M, K, float(rng.uniform(0.002, 0.08)), seed=int(rng.integers(1 << 30)),
                                   empty_rows=float(rng.choice([0.0, 0.3])))
    elif kind == "powerlaw":
        M, K = int(rng.integers(100, 3000)), int(rng.integers(50, 2000))
        m = synth.powerlaw_csr(M=M, K=K, alpha=2.1, max_deg=K, target_nnz=M * 8, seed=int(rng.integers(1 << 30)))
    elif kind == "dense_row":
        M, K = int(rng.integers(2, 600)), int(rng.integers(100, 1500))
        deg = np.zeros(M, dtype=np.int64)
        deg[int(rng.integers(M))] = K
        deg[rng.random(M) < 0.05] = 2
        m = synth.csr_from_degrees(rng, deg, K)
    else:  # empty
        M, K = int(rng.integers(1, 300)), int(rng.integers(1, 300))
        m = synth.csr_from_degrees(rng, np.zeros(M, dtype=np.int64), K)
    return synth.HostCsr(m.nrow, m.ncol, m.rowptr, m.colind, _int_values(rng, m.nnz))


def _spmm_knobs(rng, variant):
    bd = dict(block_numer=1, block_denom=int(rng.choice(DENOMS))) if rng.random() < 0.4 else {}
    tpb = dict(thread_per_block=int(rng.choice(TPB))) if rng.random() < 0.5 else {}
    if variant in ("ebpr", "rbpr"):
        gf = int(rng.integers(0, 7))  # a group never spans more than its tile
        return dict(group_factor=gf, tile_factor=int(rng.integers(gf, 7)), **tpb, **bd)
    if variant == "ebsr":
        return dict(group_factor=int(rng.integers(0, 7)), coarsen_factor=int(rng.choice([1, 2, 4, 8])),
                    thread_nz=int(rng.choice([1, 2, 3, 5, 8, 16, 32, 64])), write_policy=int(rng.integers(1, 4)),
                    **tpb, **bd)
    if variant == "rbsr":
        return dict(coarsen_factor=int(rng.choice([1, 2, 4, 8])), tile_factor=int(rng.integers(0, 7)), **tpb, **bd)
    return None


def _spmm_cases(n=180):
    rng = np.random.default_rng(SEED)
    variants = ["plain_ebpr", "plain_rbpr", "plain_ebsr", "plain_rbsr", "ebpr", "rbpr", "ebsr", "rbsr", "auto"]
    kinds = ["random", "powerlaw", "dense_row", "empty"]
    out = []
    for i in range(n):
        v = variants[i % len(variants)]
        kind = str(rng.choice(kinds, p=[0.4, 0.35, 0.2, 0.05]))
        out.append(pytest.param(i, v, kind, int(rng.choice(NS)), str(rng.choice(list(DTYPES))),
                                _spmm_knobs(rng, v), int(rng.integers(1 << 30)), id=f"spmm{i}-{v}-{kind}"))
    return out


def _rejected(e):
    return isinstance(e, _abi.SegscanError) and e.status in (_abi.EUNSUPPORTED, _abi.EINVAL)


@pytest.mark.parametrize("i,variant,kind,N,dt,knobs,seed", _spmm_cases())
def test_spmm_knob_sweep_exact(gpu, oracle, i, variant, kind, N, dt, knobs, seed):
    rng = np.random.default_rng(seed)
    m = _matrix(rng, kind)
    B = rng.integers(-3, 4, size=(m.ncol, N)).astype(np.float32)
    ref = oracle.spmm_f32(m.rowptr, m.colind, m.vals, B)
    A = m.to_device(gpu)
    try:
        C = ops.spmm(A, torch.from_numpy(B).to(gpu).to(DTYPES[dt]), variant=variant, knobs=knobs)
    except _abi.SegscanError as e:
        if _rejected(e):
            pytest.skip(f"not compiled: {e}")
        raise
    got = C.cpu().numpy()
    assert np.array_equal(got, ref), f"{variant} {knobs} {dt} N={N}: max diff {np.abs(got - ref).max()}"


def _segment_knobs(rng, variant):
    if variant == "segment_coo":
        gf = int(rng.integers(0, 7))
        return dict(group_factor=gf, tile_factor=int(rng.integers(gf, 7)))
    if variant == "segment_coo_new":
        gf = int(rng.integers(0, 7))
        return dict(group_factor=gf, tile_factor=int(rng.integers(gf, 7)), block_dim_y=int(rng.choice([1, 2, 4, 8])))
    if variant in ("segment_coo_sr", "segment_coo_noshmem_sr", "segment_coo_noshmem_lessatom_sr"):
        return dict(group_factor=int(rng.integers(0, 7)), coarsen_factor=int(rng.choice([1, 2, 4, 8])),
                    thread_nz=int(rng.choice([1, 2, 5, 16, 32])))
    if variant == "segment_coo_sorted_sr":
        return dict(coarsen_factor=int(rng.choice([1, 2, 4])), thread_nz=int(rng.choice([1, 3, 8, 16])),
                    dtile_size=int(rng.choice([1, 2, 4, 8, 16, 32])), dtile_num=int(rng.integers(1, 4)),
                    block_dim_y=int(rng.integers(1, 17)))
    return None


def _segment_cases(n=98):
    rng = np.random.default_rng(SEED + 1)
    variants = ["segment_coo", "segment_coo_new", "segment_coo_sr", "segment_coo_noshmem_sr",
                "segment_coo_noshmem_lessatom_sr", "segment_coo_sorted_sr", "auto"]
    out = []
    for i in range(n):
        v = variants[i % len(variants)]
        out.append(pytest.param(i, v, int(rng.choice(NS)), str(rng.choice(list(DTYPES))), _segment_knobs(rng, v),
                                int(rng.integers(1 << 30)), id=f"seg{i}-{v}"))
    return out


def _index(rng):
    """Sorted index with power-law runs, empty segments (also at both ends) and a dst
    longer than the last used segment."""
    S = int(rng.integers(1, 1500))
    deg = np.clip(rng.zipf(2.1, size=S), 1, int(rng.choice([50, 2000, 6000])))
    deg[rng.random(S) < 0.2] = 0
    if deg.sum() == 0:
        deg[-1] = 1
    idx = synth.segment_index_from_degrees(deg)
    return idx, S + int(rng.integers(0, 5))


@pytest.mark.parametrize("i,variant,N,dt,knobs,seed", _segment_cases())
def test_segment_knob_sweep_exact(gpu, oracle, i, variant, N, dt, knobs, seed):
    rng = np.random.default_rng(seed)
    idx, dst_len = _index(rng)
    src = rng.integers(-3, 4, size=(idx.size, N)).astype(np.float32)
    ref = oracle.segment_coo_f32(src, idx, dst_len)
    try:
        dst = ops.segment_coo(torch.from_numpy(src).to(gpu).to(DTYPES[dt]), torch.from_numpy(idx).to(gpu), dst_len,
                              variant=variant, knobs=knobs)
    except _abi.SegscanError as e:
        if _rejected(e):
            pytest.skip(f"not compiled: {e}")
        raise
    got = dst.cpu().numpy()
    assert np.array_equal(got, ref), f"{variant} {knobs} {dt} N={N}: max diff {np.abs(got - ref).max()}"


def _backward_cases(n=50):
    rng = np.random.default_rng(SEED + 2)
    combos = [("f32", "f32"), ("bf16", "bf16"), ("f16", "f16"), ("f32", "bf16"), ("f32", "f16")]
    out = []
    for i in range(n):
        knobs = {}
        if rng.random() < 0.6:
            knobs["thread_nz"] = int(rng.choice([1, 2, 3, 8, 16, 40]))
        if rng.random() < 0.4:
            knobs["thread_per_block"] = int(rng.choice([64, 128, 256]))
        if rng.random() < 0.3:
            knobs.update(block_numer=1, block_denom=int(rng.choice(DENOMS)))
        if rng.random() < 0.3:
            knobs["coarsen_factor"] = int(rng.choice([1, 2, 4]))
        xt, yt = combos[i % len(combos)]
        kind = str(rng.choice(["random", "powerlaw", "dense_row", "empty"], p=[0.35, 0.4, 0.2, 0.05]))
        out.append(pytest.param(i, kind, int(rng.choice([1, 3, 8, 16, 33, 64, 100, 128])), xt, yt, knobs or None,
                                bool(rng.random() < 0.25), int(rng.integers(1 << 30)), id=f"bwd{i}-{kind}-{xt}-{yt}"))
    return out


@pytest.mark.parametrize("i,kind,N,xt,yt,knobs,ones,seed", _backward_cases())
def test_fused_backward_knob_sweep_exact(gpu, i, kind, N, xt, yt, knobs, ones, seed):
    rng = np.random.default_rng(seed)
    m = _matrix(rng, kind)
    if ones:
        m = synth.HostCsr(m.nrow, m.ncol, m.rowptr, m.colind, None)
    dC = rng.integers(-3, 4, size=(m.nrow, N)).astype(np.float32)
    B = rng.integers(-3, 4, size=(m.ncol, N)).astype(np.float32)
    A = m.to_device(gpu)
    T = ops.csr_transpose(A)
    try:
        dB, dv = ops.spmm_backward(T, A.vals, torch.from_numpy(dC).to(gpu).to(DTYPES[xt]),
                                   torch.from_numpy(B).to(gpu).to(DTYPES[yt]), knobs=knobs)
    except _abi.SegscanError as e:
        if _rejected(e):
            pytest.skip(f"not compiled: {e}")
        raise
    vals = m.vals.astype(np.float64) if m.vals is not None else np.ones(m.nnz)
    rows = m.row_ids()
    dB_ref = np.zeros((m.ncol, N))
    np.add.at(dB_ref, m.colind, vals[:, None] * dC[rows].astype(np.float64))
    dv_ref = np.einsum("pn,pn->p", dC[rows].astype(np.float64), B[m.colind].astype(np.float64))
    assert np.array_equal(dB.cpu().numpy(), dB_ref.astype(np.float32)), f"dB {knobs}"
    assert np.array_equal(dv.cpu().numpy(), dv_ref.astype(np.float32)), f"dvals {knobs}"
