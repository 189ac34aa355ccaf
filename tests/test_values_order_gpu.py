"""Learnable values kept in the transpose's order: the forward reads entry p's value at
vals_map[p] (the transpose's inverse permutation), the fused backward reads the values
and writes dvals in its own walk order. Against the CSR-order path on the same numbers:
integer operands make every result bit-identical; the autograd wrapper's gradients
equal the CSR-order ones after the permutation."""
import numpy as np
import pytest
import torch

from segscan_amd import autograd, ops, synth

pytestmark = pytest.mark.gpu

VARIANTS = ["plain_ebpr", "plain_rbpr", "plain_ebsr", "plain_rbsr", "ebpr", "rbpr", "ebsr", "rbsr", "auto",
            ("ebsr", dict(group_factor=5, coarsen_factor=2, thread_nz=8, write_policy=2))]


@pytest.fixture(scope="module")
def mat():
    return synth.powerlaw_csr(M=6000, K=5000, max_deg=2000, target_nnz=90_000, seed=21, values="int")


def test_forward_through_vals_map_equals_csr_order(gpu, oracle, mat):
    m = mat
    A = m.to_device(gpu)
    T = ops.csr_transpose(A)
    iperm = ops.invert_permutation(T.perm)
    assert torch.equal(iperm[T.perm.long()].cpu(), torch.arange(m.nnz, dtype=torch.int32))
    vals_t = A.vals.index_select(0, T.perm.long())
    At_order = A.with_vals_map(vals_t, iperm)
    rng = np.random.default_rng(22)
    for N in (1, 16, 64):
        Bh = rng.integers(-2, 3, size=(m.ncol, N)).astype(np.float32)
        B = torch.from_numpy(Bh).to(gpu)
        ref = oracle.spmm_f32(m.rowptr, m.colind, m.vals, Bh)
        for case in VARIANTS:
            v, k = case if isinstance(case, tuple) else (case, None)
            C = ops.spmm(At_order, B, variant=v, knobs=k).cpu().numpy()
            assert np.array_equal(C, ref), (v, k, N)


@pytest.mark.parametrize("vals_in", ["at_order", "through_perm", "ones"])
def test_backward_dvals_at_order(gpu, mat, vals_in):
    m = mat
    A = m.to_device(gpu)
    T = ops.csr_transpose(A)
    rng = np.random.default_rng(23)
    N = 64
    B = torch.from_numpy(rng.integers(-2, 3, size=(m.ncol, N)).astype(np.float32)).to(gpu)
    dC = torch.from_numpy(rng.integers(-2, 3, size=(m.nrow, N)).astype(np.float32)).to(gpu)
    vals = None if vals_in == "ones" else A.vals
    vals_t = A.vals.index_select(0, T.perm.long()) if vals_in == "at_order" else None
    dB_ref, dv_ref = ops.spmm_backward(T, vals, dC, B, vals_t=vals_t)
    perm = None if vals_in == "at_order" else T.perm
    T2 = ops.CscMatrix(T.t, perm if perm is not None else torch.empty(0, dtype=torch.int32, device=gpu))
    dB, dv = ops.spmm_backward(T2 if vals_in == "at_order" else T, vals, dC, B, vals_t=vals_t, dvals_at_order=True)
    assert torch.equal(dB, dB_ref)
    assert torch.equal(dv, dv_ref.index_select(0, T.perm.long()))


def test_autograd_transpose_order_values(gpu):
    m = synth.powerlaw_csr(M=3000, K=3000, max_deg=500, target_nnz=40_000, seed=24)
    A = m.to_device(gpu)
    rng = np.random.default_rng(25)
    N = 64
    X = torch.from_numpy(rng.standard_normal((m.ncol, N)).astype(np.float32)).to(gpu)
    dC = torch.from_numpy(rng.standard_normal((m.nrow, N)).astype(np.float32)).to(gpu)
    grads = {}
    for order in ("csr", "transpose"):
        pat = autograd.SpmmPattern(A, values_order=order)
        vals = pat.to_pattern_order(A.vals).clone().requires_grad_()
        Xg = X.clone().requires_grad_()
        C = autograd.spmm(vals, Xg, pat)
        C.backward(dC)
        grads[order] = (C.detach(), vals.grad, Xg.grad, pat)
    C0, dv0, dX0, p0 = grads["csr"]
    C1, dv1, dX1, p1 = grads["transpose"]
    # the forward and dX differ only in the order carries are added (rows crossing chunk
    # boundaries); every dvals entry is one dot product, equal after the permutation
    assert torch.allclose(C0, C1, rtol=1e-5, atol=1e-5)
    assert torch.allclose(dX0, dX1, rtol=1e-5, atol=1e-5)
    assert torch.equal(dv1, dv0.index_select(0, p1.T.perm.long()))
