"""The hipSPARSE comparator computes the same product as the oracle, so the
vendor timings in bench.py compare like with like (the reference never checked
its cuSPARSE output, spmm.cuh:696-716 is unused)."""
import numpy as np
import pytest
import torch

from segscan_amd import synth

pytestmark = pytest.mark.gpu


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
