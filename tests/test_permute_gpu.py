"""The two-pass permutation plan (segscan_permute_plan_*): out[i] = x[src[i]] bit for bit
against an index gather, over sizes around its tile (16384) and bucket (8192) lengths,
random permutations and the transpose's inverse permutation of a power-law matrix, two
runs of one plan in a row; a src that is not a permutation is refused.

segscan_spmm_csr_permuted (the SpMM that reads its values through a plan) equals the
permutation into a temporary followed by segscan_spmm_csr (bit for bit on integer operands,
within 1e-6 of sum |a||b| on U(-1,1) ones): with the second pass
fused into the SpMM (workgroup rounds of 8192, 4096 and 2048 entries, several rounds per
workgroup, a partial last round, two column blocks, bf16 and fp16 B, accumulate) and without it
(launches it cannot fuse)."""
import ctypes

import numpy as np
import pytest
import torch

from segscan_amd import _abi, autograd, ops, synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 2, 5, 1000, 8191, 8192, 8193, 16383, 16384, 16385, 40000, 100003])
def test_random_permutation_exact(gpu, n):
    rng = np.random.default_rng(n)
    src = torch.from_numpy(rng.permutation(n).astype(np.int32)).to(gpu)
    plan = ops.PermutePlan(src)
    for seed in (0, 1):
        x = torch.from_numpy(np.random.default_rng(seed).standard_normal(n).astype(np.float32)).to(gpu)
        out = plan(x)
        assert torch.equal(out, x[src.long()]), (n, seed)


def test_identity_and_reverse(gpu):
    n = 50000
    x = torch.arange(n, dtype=torch.float32, device=gpu)
    ident = ops.PermutePlan(torch.arange(n, dtype=torch.int32, device=gpu))
    assert torch.equal(ident(x), x)
    rev = ops.PermutePlan(torch.arange(n - 1, -1, -1, dtype=torch.int32, device=gpu))
    assert torch.equal(rev(x), x.flip(0))


def test_transpose_order_values_into_csr_order(gpu):
    m = synth.powerlaw_csr(seed=0)  # 262144^2, 4.19M entries
    A = m.to_device(gpu)
    T = ops.csr_transpose(A)
    iperm = ops.invert_permutation(T.perm)
    vals_t = A.vals.index_select(0, T.perm.long())
    plan = ops.PermutePlan(iperm)
    out = torch.full_like(vals_t, float("nan"))
    plan(vals_t, out=out)
    assert torch.equal(out, A.vals)
    plan(vals_t * 2, out=out)
    assert torch.equal(out, A.vals * 2)


def test_not_a_permutation_refused(gpu):
    src = torch.tensor([0, 1, 1, 3], dtype=torch.int32, device=gpu)
    with pytest.raises(_abi.SegscanError) as e:
        ops.PermutePlan(src)
    assert e.value.status == _abi.EINVAL and "permutation" in str(e.value)
    src = torch.tensor([0, 4, 1, 2], dtype=torch.int32, device=gpu)
    with pytest.raises(_abi.SegscanError):
        ops.PermutePlan(src)


def test_autograd_forward_values_modes_agree(gpu):
    m = synth.powerlaw_csr(M=6000, K=5000, max_deg=2000, target_nnz=90_000, seed=31, values="int")
    A = m.to_device(gpu)
    rng = np.random.default_rng(32)
    B = torch.from_numpy(rng.integers(-2, 3, size=(m.ncol, 64)).astype(np.float32)).to(gpu)
    outs = []
    for mode in ("permute", "vals_map"):
        pat = autograd.SpmmPattern(A, values_order="transpose", forward_values=mode)
        vals = pat.to_pattern_order(A.vals).clone().requires_grad_(True)
        Bp = B.clone().requires_grad_(True)
        C = autograd.spmm(vals, Bp, pat)
        C.backward(torch.ones_like(C))
        outs.append((C.detach(), vals.grad, Bp.grad))
    for a, b in zip(outs[0], outs[1]):
        assert torch.equal(a, b)


def test_autograd_csr_order_backward_modes_agree(gpu):
    """Values in A's order: the backward through two permutation plans, through perm inside
    the walk, and "auto" (the default: both timed on the first backward, the faster kept for
    later ones) give the same gradients (integer operands: exact)."""
    m = synth.powerlaw_csr(M=6000, K=5000, max_deg=2000, target_nnz=90_000, seed=33, values="int")
    A = m.to_device(gpu)
    rng = np.random.default_rng(34)
    B = torch.from_numpy(rng.integers(-2, 3, size=(m.ncol, 64)).astype(np.float32)).to(gpu)
    dC = torch.from_numpy(rng.integers(-2, 3, size=(m.nrow, 64)).astype(np.float32)).to(gpu)
    outs = []
    for mode in ("permute", "perm", "auto"):
        pat = autograd.SpmmPattern(A, backward_values=mode)
        assert (pat.csr_permute_plans() is not None) == (mode != "perm")
        for _ in range(2 if mode == "auto" else 1):  # auto: the timed first call, then the kept route
            vals = A.vals.clone().requires_grad_(True)
            Bp = B.clone().requires_grad_(True)
            C = autograd.spmm(vals, Bp, pat)
            C.backward(dC)
            outs.append((C.detach(), vals.grad, Bp.grad))
        if mode == "auto":
            assert list(pat.backward_choice.values()) in (["permute"], ["perm"])
        else:
            assert pat.backward_choice == {}
    assert autograd.SpmmPattern(A).backward_values == "auto"
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert torch.equal(a, b)


def _spmm_permuted_case(gpu, m, N, dtype, variant, knobs, accumulate=False, seed=0):
    """Integer operands (every partial sum exact, whatever order the carries of a row that
    spans chunks are added in): bit for bit. U(-1,1) operands: within 1e-6 of
    sum_j |a_ij| |b_jk| (the carries' atomic order may round differently from run to run)."""
    A = m.to_device(gpu)
    T = ops.csr_transpose(A)
    plan = ops.PermutePlan(ops.invert_permutation(T.perm))
    rng = np.random.default_rng(seed)
    for integer in (True, False):
        if integer:
            vals_t = torch.from_numpy(rng.integers(-2, 3, size=m.nnz).astype(np.float32)).to(gpu)
            B = torch.from_numpy(rng.integers(-2, 3, size=(m.ncol, N)).astype(np.float32)).to(gpu).to(dtype)
            C0 = torch.from_numpy(rng.integers(-2, 3, size=(m.nrow, N)).astype(np.float32)).to(gpu)
        else:
            vals_t = torch.from_numpy(rng.uniform(-1, 1, size=m.nnz).astype(np.float32)).to(gpu)
            B = torch.from_numpy(rng.uniform(-1, 1, size=(m.ncol, N)).astype(np.float32)).to(gpu).to(dtype)
            C0 = torch.from_numpy(rng.uniform(-1, 1, size=(m.nrow, N)).astype(np.float32)).to(gpu)
        ref = C0.clone()
        ops.spmm(A.with_vals(plan(vals_t)), B, variant=variant, knobs=knobs, out=ref, accumulate=accumulate)
        out = torch.full_like(C0, float("nan")) if not accumulate else C0.clone()
        plan.spmm(A, vals_t, B, variant=variant, knobs=knobs, out=out, accumulate=accumulate)
        if integer:
            assert torch.equal(out, ref), (N, dtype, variant, knobs, accumulate)
        else:
            scale = ops.spmm(A.with_vals(plan(vals_t.abs())), B.abs(), variant=variant, knobs=knobs)
            err = ((out - ref).abs().max() / scale.max().clamp_min(1e-30)).item()
            assert err <= 1e-6, (err, N, dtype, variant, knobs, accumulate)
        # the plan holds no state between calls: a second call with other values
        vals2 = vals_t * -2 if integer else vals_t * -0.5
        ops.spmm(A.with_vals(plan(vals2)), B, variant=variant, knobs=knobs, out=ref)
        plan.spmm(A, vals2, B, variant=variant, knobs=knobs, out=out)
        if integer:
            assert torch.equal(out, ref)
    return ops.spmm_describe(A, N, dtype, variant, knobs)


# thread_nz x group x (thread_per_block / group) = a workgroup round's entries
_FUSED = [
    ("ebsr", dict(group_factor=5, thread_nz=32, thread_per_block=256, coarsen_factor=2, write_policy=2)),  # 8192
    ("ebsr", dict(group_factor=5, thread_nz=16, thread_per_block=256, coarsen_factor=2, write_policy=2)),  # 4096
    ("ebsr", dict(group_factor=4, thread_nz=16, thread_per_block=128, coarsen_factor=4, write_policy=2)),  # 2048
    ("ebsr", dict(group_factor=5, thread_nz=32, thread_per_block=256, coarsen_factor=2, write_policy=3)),  # lessatom
    ("ebsr", dict(group_factor=5, thread_nz=32, thread_per_block=256, coarsen_factor=2, write_policy=2,
                  block_numer=1, block_denom=1000)),  # a few workgroups, several rounds each
]
_UNFUSED = [
    ("ebsr", dict(group_factor=5, thread_nz=4, thread_per_block=256, coarsen_factor=2)),  # 1024-entry rounds
    ("ebsr", dict(group_factor=5, thread_nz=32, thread_per_block=192, coarsen_factor=2)),  # 6144: not a divisor
    ("ebsr", dict(group_factor=2, thread_nz=32, thread_per_block=64, coarsen_factor=8)),  # 2048, 4-lane groups
    ("rbsr", None),
    ("ebpr", None),
]


def _fuses(info) -> bool:
    """The library's rule: ebsr with groups of >= 8 lanes, a workgroup round's entries
    dividing 8192, >= 2048."""
    span = info["chunk"] * (info["thread_per_block"] // info["group_size"])
    return (info["variant"] == _abi.SPMM_EBSR and info["group_size"] >= 8 and span >= 2048
            and 8192 % span == 0)


@pytest.mark.parametrize("nnz", [90_000, 70_001])
@pytest.mark.parametrize("case", range(len(_FUSED) + len(_UNFUSED) + 1))
def test_spmm_permuted_equals_permute_then_spmm(gpu, nnz, case):
    variant, knobs = (_FUSED + _UNFUSED + [("auto", None)])[case]
    m = synth.powerlaw_csr(M=6000, K=5000, max_deg=2000, target_nnz=nnz, seed=41 + nnz % 7)
    for N, dtype in ((64, torch.float32), (128, torch.float32), (64, torch.bfloat16), (64, torch.float16),
                     (16, torch.float32)):
        info = _spmm_permuted_case(gpu, m, N, dtype, variant, knobs, seed=N)
        if case < len(_FUSED):
            assert _fuses(info), info
        elif case < len(_FUSED) + len(_UNFUSED):
            assert not _fuses(info), info
    _spmm_permuted_case(gpu, m, 64, torch.float32, variant, knobs, accumulate=True, seed=5)


def test_spmm_permuted_cases_cover_partial_rounds():
    """Host check of the cases above: some fused case ends in a workgroup round whose last
    groups have no chunk (nchunks not a multiple of the groups per workgroup)."""
    partial = False
    for nnz in (90_000, 70_001):
        m = synth.powerlaw_csr(M=6000, K=5000, max_deg=2000, target_nnz=nnz, seed=41 + nnz % 7)
        for variant, knobs in _FUSED:
            info = ops.spmm_describe(m.to_device("cpu"), 64, torch.float32, variant, knobs)
            gpb = info["thread_per_block"] // info["group_size"]
            partial |= info["nchunks"] % gpb != 0
    assert partial


def test_spmm_permuted_headline_points(gpu):
    """The power-law matrix (4.19M entries) at the tuned fp32 (8192-entry rounds) and bf16
    (4096-entry rounds) points of the tuning table, N = 64."""
    m = synth.powerlaw_csr(seed=0)
    fp32 = dict(group_factor=5, thread_nz=32, thread_per_block=256, coarsen_factor=2, write_policy=2,
                block_numer=1, block_denom=512)
    bf16 = dict(group_factor=5, thread_nz=16, thread_per_block=256, coarsen_factor=2, write_policy=2,
                block_numer=1, block_denom=256)
    _spmm_permuted_case(gpu, m, 64, torch.float32, "ebsr", fp32)
    _spmm_permuted_case(gpu, m, 64, torch.bfloat16, "ebsr", bf16)


def test_spmm_permuted_refusals(gpu):
    m = synth.powerlaw_csr(M=600, K=500, max_deg=200, target_nnz=5000, seed=3)
    A = m.to_device(gpu)
    plan = ops.PermutePlan(torch.arange(m.nnz - 1, -1, -1, dtype=torch.int32, device=gpu))
    B = torch.ones((m.ncol, 8), device=gpu)
    lib = _abi.load()
    desc = A.desc()  # A's own values set: refused
    C = torch.empty((m.nrow, 8), device=gpu)
    ws = torch.empty(1 << 20, dtype=torch.uint8, device=gpu)
    v = torch.ones(m.nnz, device=gpu)
    st = lib.segscan_spmm_csr_permuted(ctypes.byref(desc), plan._h, v.data_ptr(), 8, B.data_ptr(), 8, _abi.F32,
                                       C.data_ptr(), 8, _abi.SPMM_AUTO, None, 0, ws.data_ptr(), ws.numel(), None)
    assert st == _abi.EINVAL and b"must be NULL" in lib.segscan_last_error()
    small = ops.PermutePlan(torch.arange(m.nnz - 1, dtype=torch.int32, device=gpu))
    with pytest.raises(ValueError):
        small.spmm(A, v[:-1], B)
    desc.vals = None
    st = lib.segscan_spmm_csr_permuted(ctypes.byref(desc), small._h, v.data_ptr(), 8, B.data_ptr(), 8, _abi.F32,
                                       C.data_ptr(), 8, _abi.SPMM_AUTO, None, 0, ws.data_ptr(), ws.numel(), None)
    assert st == _abi.EINVAL and b"entries" in lib.segscan_last_error()
