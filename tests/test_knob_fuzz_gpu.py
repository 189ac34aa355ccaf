"""Seeded knob-space sweep on the GPU: SpMM (every variant), segment_coo (the six
reference launchers and AUTO) and the fused backward, each with knobs, shapes and
storage types drawn from a fixed seed, checked bit for bit against the CPU oracle.

Every operand is a small integer (A's values and the dense rows in {-3..3}), so each
partial sum is exact in fp32 in any order and in bf16 / fp16 storage; the GPU result
must then equal the oracle exactly, whatever the lane layout, chunking or write
policy. The support contract is the host-side plan (the describe / workspace-size
calls, which run without a GPU): a case the plan accepts must run and match exactly;
a case it rejects must be rejected by the call with the same status and message,
before anything is launched. The CPU tests pin the plan itself: every drawn case is
supported except the fused backward's G x CF < N ones, and each knob limit the
planners enforce is hit by a case of its own. The cases are fixed by the seed, one
test item each, so a failure names its case.
"""
import numpy as np
import pytest
import torch

from segscan_amd import _abi, ops, synth

gpu_only = pytest.mark.gpu

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


def _plan(fn):
    """The host plan's verdict: None when accepted, else (status, message)."""
    try:
        fn()
    except _abi.SegscanError as e:
        return e.status, str(e).split(": ", 1)[1]
    return None


def _host_csr(nrow, ncol, nnz):
    """Descriptor with host arrays: the plans read sizes only."""
    return ops.CsrMatrix.from_numpy(nrow, ncol, np.zeros(nrow + 1, np.int32), np.zeros(nnz, np.int32), None,
                                    device="cpu")


def _spmm_plan(m, N, dt, variant, knobs):
    return _plan(lambda: ops.spmm_describe(_host_csr(m.nrow, m.ncol, m.nnz), N, DTYPES[dt], variant, knobs))


def _segment_plan(E, N, dst_len, dt, variant, knobs):
    return _plan(lambda: ops.segment_describe(E, N, dst_len, DTYPES[dt], variant, knobs))


def _backward_plan(m, N, xt, knobs):
    T = ops.CscMatrix(_host_csr(m.ncol, m.nrow, m.nnz), torch.zeros(m.nnz, dtype=torch.int32))
    return _plan(lambda: ops.spmm_backward_workspace(T, N, DTYPES[xt], knobs, device="cpu"))


def _expect(plan, call):
    """Runs call; a rejected plan must be rejected the same way, with nothing launched."""
    if plan is None:
        return call()
    with pytest.raises(_abi.SegscanError) as exc:
        call()
    assert (exc.value.status, str(exc.value).split(": ", 1)[1]) == plan
    return None


@gpu_only
@pytest.mark.parametrize("i,variant,kind,N,dt,knobs,seed", _spmm_cases())
def test_spmm_knob_sweep_exact(gpu, oracle, i, variant, kind, N, dt, knobs, seed):
    rng = np.random.default_rng(seed)
    m = _matrix(rng, kind)
    B = rng.integers(-3, 4, size=(m.ncol, N)).astype(np.float32)
    ref = oracle.spmm_f32(m.rowptr, m.colind, m.vals, B)
    A = m.to_device(gpu)
    C = _expect(_spmm_plan(m, N, dt, variant, knobs),
                lambda: ops.spmm(A, torch.from_numpy(B).to(gpu).to(DTYPES[dt]), variant=variant, knobs=knobs))
    if C is None:
        return
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


@gpu_only
@pytest.mark.parametrize("i,variant,N,dt,knobs,seed", _segment_cases())
def test_segment_knob_sweep_exact(gpu, oracle, i, variant, N, dt, knobs, seed):
    rng = np.random.default_rng(seed)
    idx, dst_len = _index(rng)
    src = rng.integers(-3, 4, size=(idx.size, N)).astype(np.float32)
    ref = oracle.segment_coo_f32(src, idx, dst_len)
    dst = _expect(_segment_plan(idx.size, N, dst_len, dt, variant, knobs),
                  lambda: ops.segment_coo(torch.from_numpy(src).to(gpu).to(DTYPES[dt]), torch.from_numpy(idx).to(gpu),
                                          dst_len, variant=variant, knobs=knobs))
    if dst is None:
        return
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
        # dc_map: A's row r reads dC[dc_map[r]] from a dC of R rows (0 = no map)
        R = int(rng.integers(1, 40)) if rng.random() < 0.3 else 0
        out.append(pytest.param(i, kind, int(rng.choice([1, 3, 8, 16, 33, 64, 100, 128])), xt, yt, knobs or None,
                                bool(rng.random() < 0.25), R, int(rng.integers(1 << 30)),
                                id=f"bwd{i}-{kind}-{xt}-{yt}" + (f"-map{R}" if R else "")))
    return out


@gpu_only
@pytest.mark.parametrize("i,kind,N,xt,yt,knobs,ones,R,seed", _backward_cases())
def test_fused_backward_knob_sweep_exact(gpu, i, kind, N, xt, yt, knobs, ones, R, seed):
    rng = np.random.default_rng(seed)
    m = _matrix(rng, kind)
    if ones:
        m = synth.HostCsr(m.nrow, m.ncol, m.rowptr, m.colind, None)
    dc_map = rng.integers(0, R, size=m.nrow).astype(np.int32) if R else None
    dC = rng.integers(-3, 4, size=(R if R else m.nrow, N)).astype(np.float32)
    B = rng.integers(-3, 4, size=(m.ncol, N)).astype(np.float32)
    A = m.to_device(gpu)
    T = ops.csr_transpose(A)
    dm = torch.from_numpy(dc_map).to(gpu) if R else None
    res = _expect(_backward_plan(m, N, xt, knobs),
                  lambda: ops.spmm_backward(T, A.vals, torch.from_numpy(dC).to(gpu).to(DTYPES[xt]),
                                            torch.from_numpy(B).to(gpu).to(DTYPES[yt]), knobs=knobs, dc_map=dm))
    if res is None:
        return
    dB, dv = res
    vals = m.vals.astype(np.float64) if m.vals is not None else np.ones(m.nnz)
    rows = m.row_ids()
    dC_rows = dC[dc_map[rows] if R else rows].astype(np.float64)
    dB_ref = np.zeros((m.ncol, N))
    np.add.at(dB_ref, m.colind, vals[:, None] * dC_rows)
    dv_ref = np.einsum("pn,pn->p", dC_rows, B[m.colind].astype(np.float64))
    assert np.array_equal(dB.cpu().numpy(), dB_ref.astype(np.float32)), f"dB {knobs}"
    assert np.array_equal(dv.cpu().numpy(), dv_ref.astype(np.float32)), f"dvals {knobs}"


# ---- the plans themselves (CPU) ---------------------------------------------------------

def test_drawn_cases_supported_except_narrow_backward_groups():
    """Every drawn SpMM and segment case plans; a drawn backward case is refused exactly
    when its coarsen_factor leaves 64 lanes (the widest group) short of N columns
    (segscan_abi.cpp plan_bwd: group_size x coarsen_factor >= N)."""
    for prm in _spmm_cases():
        i, v, kind, N, dt, knobs, seed = prm.values
        m = _matrix(np.random.default_rng(seed), kind)
        assert _spmm_plan(m, N, dt, v, knobs) is None, prm.id
    for prm in _segment_cases():
        i, v, N, dt, knobs, seed = prm.values
        idx, dst_len = _index(np.random.default_rng(seed))
        assert _segment_plan(idx.size, N, dst_len, dt, v, knobs) is None, prm.id
    refused = 0
    for prm in _backward_cases():
        i, kind, N, xt, yt, knobs, ones, R, seed = prm.values
        m = _matrix(np.random.default_rng(seed), kind)
        cf = (knobs or {}).get("coarsen_factor")
        plan = _backward_plan(m, N, xt, knobs)
        if cf and 64 * cf < N:
            refused += 1
            assert plan is not None and plan[0] == _abi.EUNSUPPORTED and "group_size x coarsen_factor >= N" in plan[1]
        else:
            assert plan is None, (prm.id, plan)
    assert refused >= 1  # the seed draws some


M_SMALL = synth.HostCsr(4, 5, np.array([0, 1, 1, 3, 4], np.int32), np.array([0, 1, 4, 2], np.int32), None)

# (op, variant, N, dtype, knobs, message fragment): one case per limit the planners enforce
UNSUPPORTED = [
    ("spmm", "ebsr", 64, "f32", dict(group_factor=7), "groups are at most 64 lanes"),
    ("spmm", "ebpr", 64, "f32", dict(group_factor=3, tile_factor=2), "group_size 8 exceeds tile_size 4"),
    ("spmm", "rbpr", 64, "f32", dict(tile_factor=7), "tiles are at most one wave"),
    ("spmm", "ebpr", 64, "f32", dict(tile_factor=9), "tile_factor 9 too large"),
    ("spmm", "ebpr", 64, "f32", dict(coarsen_factor=8), "coarsen_factor 8 needs a 16-bit dense operand"),
    ("spmm", "ebsr", 64, "f32", dict(coarsen_factor=3), "coarsen_factor 3 not in {1,2,4,8}"),
    ("spmm", "ebsr", 64, "f32", dict(thread_per_block=2048), "thread_per_block 2048 outside [1, 1024]"),
    ("spmm", "ebsr", 64, "f32", dict(thread_nz=(1 << 20) + 1), "thread_nz 1048577 outside [1, 2^20]"),
    ("spmm", "rbsr", 3, "f32", dict(coarsen_factor=1, tile_factor=2, thread_per_block=1024),
     "workgroup of 1026 lanes exceeds 1024"),
    ("segment", "segment_coo", 64, "f32", dict(coarsen_factor=8), "coarsen_factor 8 needs a 16-bit src"),
    ("segment", "segment_coo", 64, "bf16", dict(group_factor=4, tile_factor=3), "group_size 16 exceeds tile_size 8"),
    ("segment", "segment_coo_sr", 64, "f32", dict(group_factor=7), "groups are at most 64 lanes"),
    ("segment", "segment_coo_sorted_sr", 64, "f32", dict(dtile_size=3), "must be a power of two"),
    ("segment", "segment_coo_sorted_sr", 200, "f32", dict(dtile_size=32, dtile_num=3, coarsen_factor=1),
     "= 96 lanes exceeds a wave"),
    ("backward", None, 100, "f32", dict(coarsen_factor=1), "group_size x coarsen_factor >= N (64 x 1 < 100)"),
    ("backward", None, 64, "f32", dict(group_factor=7), "groups are at most 64 lanes"),
    ("backward", None, 64, "f32", dict(thread_per_block=512), "the fused backward takes at most 256"),
    ("backward", None, 64, "f32", dict(coarsen_factor=16), "coarsen_factor 16 not in {1,2,4,8}"),
]


@pytest.mark.parametrize("op,variant,N,dt,knobs,frag", UNSUPPORTED, ids=lambda x: None)
def test_plan_refuses_each_limit(op, variant, N, dt, knobs, frag):
    if op == "spmm":
        plan = _spmm_plan(M_SMALL, N, dt, variant, knobs)
    elif op == "segment":
        plan = _segment_plan(100, N, 10, dt, variant, knobs)
    else:
        plan = _backward_plan(M_SMALL, N, dt, knobs)
    assert plan is not None and plan[0] == _abi.EUNSUPPORTED and frag in plan[1], plan


@gpu_only
@pytest.mark.parametrize("op,variant,N,dt,knobs,frag", UNSUPPORTED, ids=lambda x: None)
def test_call_refuses_each_limit(gpu, op, variant, N, dt, knobs, frag):
    """The calls refuse what the plans refuse, with the plan's message, launching nothing."""
    rng = np.random.default_rng(1)
    A = M_SMALL.to_device(gpu)
    if op == "spmm":
        B = torch.ones((A.ncol, N), device=gpu).to(DTYPES[dt])
        _expect(_spmm_plan(M_SMALL, N, dt, variant, knobs), lambda: ops.spmm(A, B, variant=variant, knobs=knobs))
    elif op == "segment":
        idx = torch.from_numpy(np.sort(rng.integers(0, 10, 100)).astype(np.int32)).to(gpu)
        src = torch.ones((100, N), device=gpu).to(DTYPES[dt])
        _expect(_segment_plan(100, N, 10, dt, variant, knobs),
                lambda: ops.segment_coo(src, idx, 10, variant=variant, knobs=knobs))
    else:
        T = ops.csr_transpose(A)
        dC = torch.ones((A.nrow, N), device=gpu)
        B = torch.ones((A.ncol, N), device=gpu)
        _expect(_backward_plan(M_SMALL, N, dt, knobs), lambda: ops.spmm_backward(T, None, dC, B, knobs=knobs))
    torch.cuda.synchronize()
