"""The two-pass permutation plan (segscan_permute_plan_*): out[i] = x[src[i]] bit for bit
against an index gather, over sizes around its tile (16384) and bucket (8192) lengths,
random permutations and the transpose's inverse permutation of a power-law matrix, two
runs of one plan in a row; a src that is not a permutation is refused."""
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
    """Values in A's order: the backward through two permutation plans (the default) and
    through perm inside the walk give the same gradients (integer operands: exact)."""
    m = synth.powerlaw_csr(M=6000, K=5000, max_deg=2000, target_nnz=90_000, seed=33, values="int")
    A = m.to_device(gpu)
    rng = np.random.default_rng(34)
    B = torch.from_numpy(rng.integers(-2, 3, size=(m.ncol, 64)).astype(np.float32)).to(gpu)
    dC = torch.from_numpy(rng.integers(-2, 3, size=(m.nrow, 64)).astype(np.float32)).to(gpu)
    outs = []
    for mode in ("permute", "perm"):
        pat = autograd.SpmmPattern(A, backward_values=mode)
        assert (pat.csr_permute_plans() is not None) == (mode == "permute")
        vals = A.vals.clone().requires_grad_(True)
        Bp = B.clone().requires_grad_(True)
        C = autograd.spmm(vals, Bp, pat)
        C.backward(dC)
        outs.append((C.detach(), vals.grad, Bp.grad))
    for a, b in zip(outs[0], outs[1]):
        assert torch.equal(a, b)
