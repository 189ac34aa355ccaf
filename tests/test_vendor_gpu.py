"""The hipSPARSE comparator computes the same product as the oracle, so the
vendor timings in bench.py compare like with like (the reference never checked
its cuSPARSE output, spmm.cuh:696-716 is unused)."""
import numpy as np
import pytest
import torch

from segscan_amd import synth


@pytest.mark.gpu
@pytest.mark.parametrize("alg", ["default", "csr_alg1", "csr_alg2", "csr_alg3"])
def test_hipsparse_matches_oracle(gpu, oracle, alg):
    from segscan_amd.vendor import VendorSpmm
    m = synth.powerlaw_csr(M=3000, K=2500, max_deg=2000, target_nnz=40000, seed=11)
    rng = np.random.default_rng(11)
    B = rng.uniform(-1, 1, size=(m.ncol, 64)).astype(np.float32)
    A = m.to_device(gpu)
    Bd = torch.from_numpy(B).to(gpu)
    C = torch.zeros((m.nrow, 64), device=gpu)
    try:
        VendorSpmm(A, Bd, C, alg)()
    except RuntimeError as exc:
        pytest.skip(f"hipSPARSE algorithm {alg} not available: {exc}")
    torch.cuda.synchronize()
    C64, D = oracle.spmm_f64(m.rowptr, m.colind, m.vals, B)
    assert oracle.normalized_error(C.cpu().numpy(), C64, D) <= 1e-6


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16])
def test_hipsparse_refuses_16bit_operands_on_host(dtype):
    """The vendor SpMM has no fp32-A / 16-bit-B form (it throws instead of returning
    a status), so the comparator refuses it before any library call."""
    from segscan_amd.ops import CsrMatrix
    from segscan_amd.vendor import VendorSpmm
    A = CsrMatrix(2, 2, torch.tensor([0, 1, 2], dtype=torch.int32), torch.tensor([0, 1], dtype=torch.int32),
                  torch.ones(2))
    with pytest.raises(ValueError, match="fp32"):
        VendorSpmm(A, torch.zeros((2, 4), dtype=dtype), torch.zeros((2, 4)))
