"""The reference-named launchers (Python and C++) against the oracle, with the
reference's own conventions: atomic variants add into a caller-zeroed output,
store variants overwrite it; inputs as the reference harness builds them
(values and B all ones) and integer-valued random inputs."""
import subprocess
from pathlib import Path

import numpy as np
import pytest
import torch

from segscan_amd import reference_api as ref
from segscan_amd import synth

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent


def test_python_reference_names(gpu, oracle):
    m = synth.powerlaw_csr(M=2000, K=1500, max_deg=1000, target_nnz=20000, seed=8, values="int")
    rng = np.random.default_rng(8)
    N = 64
    Bh = rng.integers(-2, 3, size=(m.ncol, N)).astype(np.float32)
    A = m.to_device(gpu)
    B = torch.from_numpy(Bh).to(gpu)
    expect = oracle.spmm_f32(m.rowptr, m.colind, m.vals, Bh)
    calls = [
        lambda C: ref.csrspmm_parreduce_nnzbalance(A, N, B, C),
        lambda C: ref.csrspmm_parreduce_rowbalance(A, N, B, C),
        lambda C: ref.csrspmm_rowcaching_nnzbalance(A, N, B, C),
        lambda C: ref.csrspmm_seqreduce_rowbalance(A, N, B, C),
        lambda C: ref.csrspmm_parreduce_nnzbalance_cg(A, N, B, C, 5, 256, 5, 1, 16),
        lambda C: ref.csrspmm_parreduce_rowbalance_cg(A, N, B, C, 3, 256, 3, 1, 2),
        lambda C: ref.csrspmm_rowcaching_nnzbalance_cg(A, N, B, C, 4, 256, 4, 1, 1, 16),
        lambda C: ref.csrspmm_seqreduce_rowbalance_cg(A, N, B, C, 4, 256, 5, 1, 1),
    ]
    for i, call in enumerate(calls):
        C = torch.zeros(m.nrow * N, device=gpu)  # flat buffer, as the reference's RamArray
        call(C)
        assert np.array_equal(C.view(m.nrow, N).cpu().numpy(), expect), i
    # atomic variants accumulate: a second call doubles the result
    C = torch.zeros(m.nrow * N, device=gpu)
    ref.csrspmm_parreduce_nnzbalance_cg(A, N, B, C, 5, 256, 5, 1, 16)
    ref.csrspmm_parreduce_nnzbalance_cg(A, N, B, C, 5, 256, 5, 1, 16)
    assert np.array_equal(C.view(m.nrow, N).cpu().numpy(), 2 * expect)


def test_python_segment_reference_names(gpu, oracle):
    idx, _ = oracle.generate_index(100, 1, 100, 5000, 0.5)
    N = 32
    src_h = np.random.default_rng(0).uniform(0, 1, size=(5000, N)).astype(np.float32)
    expect = oracle.segment_coo_f32(src_h, idx, 100)
    src = torch.from_numpy(src_h).to(gpu)
    index = torch.from_numpy(idx).to(gpu)
    calls = [
        lambda d: ref.segment_coo(src, index, 5000, N, d, 5, 256, 5, 1, 16),
        lambda d: ref.segment_coo_new(src, index, 5000, N, d, 5, 256, 5, 2),
        lambda d: ref.segment_coo_sr(src, index, 5000, N, d, 4, 256, 4, 1, 1, 16),
        lambda d: ref.segment_coo_noshmem_sr(src, index, 5000, N, d, 4, 256, 4, 2, 1, 16),
        lambda d: ref.segment_coo_noshmem_lessatom_sr(src, index, 5000, N, d, 4, 256, 4, 1, 1, 16),
        lambda d: ref.segment_coo_sorted_sr(src, index, 5000, N, d, 2, 4, 8, 2, 16),
    ]
    for i, call in enumerate(calls):
        dst = torch.zeros(100 * N, device=gpu)
        call(dst)
        got = dst.view(100, N).cpu().numpy()
        assert oracle.check_result(got, expect) == 0, i  # reference rule: |d - ref| <= 1e-2 |ref|
        np.testing.assert_allclose(got, expect, rtol=1e-5, atol=1e-5)


def test_cpp_harness():
    exe = ROOT / "tests" / "cpp" / "spmm_check"
    if not exe.exists():
        pytest.fail("tests/cpp/spmm_check is not built (run __graft_entry__.build())")
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(res.stdout[-4000:])
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    assert "Not Passed!" not in res.stdout
    assert res.stdout.count("Passed!") >= 50
