"""The allocation-free SpMM entry point (segscan_spmm_csr_noalloc), called through the
C ABI: its results equal segscan_spmm_csr's with an explicit workspace, on two calls in
a row (the second with a smaller input, whose output must not keep anything of the
first), for the plans that use scratch (ebsr owner carries, ebpr tile rows) and AUTO;
a plan that needs more scratch than the static array holds is refused on the host."""
import ctypes
from ctypes import byref

import numpy as np
import pytest
import torch

from segscan_amd import _abi, ops, synth

SCRATCH_BYTES = 16 << 20

CASES = [
    ("ebsr", dict(group_factor=4, coarsen_factor=4, thread_nz=4, write_policy=2)),  # owner: carries in scratch
    ("ebsr", dict(group_factor=5, coarsen_factor=2, thread_nz=2, write_policy=2, block_numer=1, block_denom=64)),
    ("ebpr", dict(group_factor=4, tile_factor=5)),                                   # tile rows in scratch
    ("auto", None),
]


def _noalloc(lib, A, B, C, variant, knobs, stream=None):
    v = _abi.SPMM_VARIANTS[variant]
    kp = None if knobs is None else byref(_abi.Knobs.make(**knobs))
    st = lib.segscan_spmm_csr_noalloc(byref(A.desc()), B.size(1), ctypes.c_void_p(B.data_ptr()), B.stride(0),
                                      _abi.F32 if B.dtype == torch.float32 else _abi.BF16,
                                      ctypes.c_void_p(C.data_ptr()), C.stride(0), v, kp, 0,
                                      ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    _abi.check(st, "spmm_csr_noalloc")


def test_noalloc_refuses_oversized_plans_on_host():
    """Planning happens before any HIP call: a plan needing more than the static scratch
    is refused with SEGSCAN_EWORKSPACE (fake device pointers are never touched)."""
    lib = _abi.load()
    d = _abi.CsrDesc()
    d.nrow, d.ncol, d.nnz = 1_000_000, 1_000_000, 10_000_000
    d.rowptr, d.colind, d.vals = 0x1000, 0x2000, None
    kn = _abi.Knobs.make(group_factor=4, coarsen_factor=4, thread_nz=1, write_policy=2)
    need = ctypes.c_size_t(0)
    assert lib.segscan_spmm_workspace_size(byref(d), 64, _abi.F32, _abi.SPMM_VARIANTS["ebsr"], byref(kn),
                                           byref(need)) == _abi.OK
    assert need.value > SCRATCH_BYTES
    st = lib.segscan_spmm_csr_noalloc(byref(d), 64, 0x10000, 64, _abi.F32, 0x20000, 64, _abi.SPMM_VARIANTS["ebsr"],
                                      byref(kn), 0, None)
    assert st == _abi.EWORKSPACE
    assert "scratch" in lib.segscan_last_error().decode()
    # argument errors come first, as in segscan_spmm_csr
    assert lib.segscan_spmm_csr_noalloc(None, 64, 0x10000, 64, _abi.F32, 0x20000, 64, 0, None, 0, None) == _abi.EINVAL


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0] + ("" if not c[1] else "-" + "-".join(
    f"{k[:2]}{v}" for k, v in c[1].items())))
def test_noalloc_matches_workspace_entry_point(gpu, oracle, case):
    variant, knobs = case
    lib = _abi.load()
    rng = np.random.default_rng(17)
    N = 64
    mats = [synth.powerlaw_csr(M=20000, K=15000, max_deg=3000, target_nnz=200_000, seed=3, values="int"),
            synth.powerlaw_csr(M=3000, K=2000, max_deg=500, target_nnz=20_000, seed=4, values="int")]
    for call, m in enumerate(mats):
        print(f"noalloc {variant} {knobs} call {call}: {m.nrow}x{m.ncol} nnz={m.nnz}", flush=True)
        A = m.to_device(gpu)
        need = ops.spmm_describe(A, N, torch.float32, variant, knobs)["workspace_bytes"]
        assert need <= SCRATCH_BYTES
        Bh = rng.integers(-2, 3, size=(m.ncol, N)).astype(np.float32)
        B = torch.from_numpy(Bh).to(gpu)
        C_ref = ops.spmm(A, B, variant=variant, knobs=knobs)
        C = torch.full((m.nrow, N), float("nan"), dtype=torch.float32, device=gpu)
        _noalloc(lib, A, B, C, variant, knobs)
        torch.cuda.synchronize()
        got = C.cpu().numpy()
        assert np.array_equal(got, C_ref.cpu().numpy()), (variant, knobs, call)
        assert np.array_equal(got, oracle.spmm_f32(m.rowptr, m.colind, m.vals, Bh)), (variant, knobs, call)
        # U(-1, 1) values: within the float64 bound, and equal to the workspace entry point
        # up to the order of the carry additions
        Bf = rng.uniform(-1, 1, size=(m.ncol, N)).astype(np.float32)
        Af = synth.HostCsr(m.nrow, m.ncol, m.rowptr, m.colind,
                           rng.uniform(-1, 1, size=m.nnz).astype(np.float32)).to_device(gpu)
        Bd = torch.from_numpy(Bf).to(gpu)
        C.fill_(float("nan"))
        _noalloc(lib, Af, Bd, C, variant, knobs)
        C_ws = ops.spmm(Af, Bd, variant=variant, knobs=knobs)
        torch.cuda.synchronize()
        C64, D = oracle.spmm_f64(m.rowptr, m.colind, Af.vals.cpu().numpy(), Bf)
        got, got_ws = C.cpu().numpy(), C_ws.cpu().numpy()
        assert oracle.normalized_error(got, C64, D) <= 1e-6
        assert oracle.normalized_error(got_ws, C64, D) <= 1e-6
        # atomic variants add in no fixed order: the two results agree within the sum of
        # their bounds, 2e-6 x sum |a||b| per element (a fixed tolerance fails on long rows)
        assert np.all(np.abs(got.astype(np.float64) - got_ws) <= 2e-6 * D), (variant, knobs, call)
