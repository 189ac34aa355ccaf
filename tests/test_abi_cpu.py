"""CPU-side checks of the C ABI: the library loads without a GPU, exports every
function include/segscan.h declares, and rejects bad arguments on the host
before any HIP call (every call here fails validation, so none reaches the
device)."""
import ctypes
import re
from ctypes import byref, c_size_t
from pathlib import Path

import pytest

from segscan_amd import _abi

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "segscan.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(segscan_[a-z_0-9]+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def lib():
    return _abi.load()


def test_header_declares_functions():
    names = declared_functions()
    assert "segscan_spmm_csr" in names and "segscan_segment_coo" in names
    assert len(names) >= 16


def test_library_exports_every_declared_symbol(lib):
    for name in declared_functions():
        assert hasattr(lib, name), f"{name} declared in segscan.h but not exported"
    # and the binding knows every declared function
    assert set(declared_functions()) == set(_abi.EXPORTED)


def test_version_and_status_strings(lib):
    assert lib.segscan_abi_version() == 3
    for s in range(5):
        assert lib.segscan_status_string(s).decode().startswith("SEGSCAN_")


def _desc(nrow=10, ncol=10, nnz=20, rowptr=1, colind=1, vals=None, max_row=0):
    d = _abi.CsrDesc()
    d.nrow, d.ncol, d.nnz = nrow, ncol, nnz
    d.rowptr, d.colind, d.vals = rowptr, colind, vals
    d.max_row_nnz = max_row
    return d


def _spmm(lib, desc, N=8, B=0x1000, ldb=8, C=0x2000, ldc=8, variant=_abi.SPMM_EBSR, knobs=None, flags=0,
          ws=None, ws_bytes=0, dtype=_abi.F32):
    kp = None if knobs is None else byref(_abi.Knobs.make(**knobs))
    return lib.segscan_spmm_csr(None if desc is None else byref(desc), N, B, ldb, dtype, C, ldc, variant, kp, flags,
                                ws, ws_bytes, None)


@pytest.mark.parametrize("case,expect", [
    (dict(desc=None), _abi.EINVAL),
    (dict(desc=_desc(nrow=-1)), _abi.EINVAL),
    (dict(desc=_desc(rowptr=None)), _abi.EINVAL),
    (dict(desc=_desc(colind=None)), _abi.EINVAL),
    (dict(desc=_desc(nrow=0, nnz=5)), _abi.EINVAL),
    (dict(desc=_desc(nnz=2**31)), _abi.EINVAL),
    (dict(desc=_desc(max_row=21)), _abi.EINVAL),
    (dict(desc=_desc(max_row=-1)), _abi.EINVAL),
    (dict(desc=_desc(), N=-1), _abi.EINVAL),
    (dict(desc=_desc(), B=None), _abi.EINVAL),
    (dict(desc=_desc(), C=None), _abi.EINVAL),
    (dict(desc=_desc(), ldb=4), _abi.EINVAL),
    (dict(desc=_desc(), ldc=7), _abi.EINVAL),
    (dict(desc=_desc(), ldb=2**31), _abi.EUNSUPPORTED),  # row offsets are 32-bit row x 32-bit stride
    (dict(desc=_desc(), C=0x2002), _abi.EINVAL),
    (dict(desc=_desc(), dtype=7), _abi.EINVAL),
    (dict(desc=_desc(), flags=4), _abi.EINVAL),
    (dict(desc=_desc(), variant=42), _abi.EINVAL),
    (dict(desc=_desc(), knobs=dict(group_factor=7)), _abi.EUNSUPPORTED),
    (dict(desc=_desc(), knobs=dict(thread_per_block=2048)), _abi.EUNSUPPORTED),
    (dict(desc=_desc(), knobs=dict(coarsen_factor=3)), _abi.EUNSUPPORTED),
    (dict(desc=_desc(), knobs=dict(block_numer=-1, block_denom=2)), _abi.EINVAL),
    (dict(desc=_desc(), knobs=dict(write_policy=9)), _abi.EINVAL),
    (dict(desc=_desc(), variant=_abi.SPMM_EBPR, knobs=dict(group_factor=6, tile_factor=5)), _abi.EUNSUPPORTED),
    (dict(desc=_desc(), variant=_abi.SPMM_EBPR, knobs=dict(tile_factor=7)), _abi.EUNSUPPORTED),
    (dict(desc=_desc(), variant=_abi.SPMM_EBPR, knobs=dict(coarsen_factor=8)), _abi.EUNSUPPORTED),
    (dict(desc=_desc(), variant=_abi.SPMM_RBSR, knobs=dict(tile_factor=9)), _abi.EUNSUPPORTED),
    # owner policy needs a workspace
    (dict(desc=_desc(), knobs=dict(write_policy=2)), _abi.EWORKSPACE),
    (dict(desc=_desc(), knobs=dict(write_policy=2), ws=0x3000, ws_bytes=1), _abi.EWORKSPACE),
])
def test_spmm_validation(lib, case, expect):
    st = _spmm(lib, **case)
    assert st == expect, lib.segscan_last_error().decode()
    assert lib.segscan_last_error().decode() != ""


def test_segment_validation(lib):
    k = None
    f = lib.segscan_segment_coo
    assert f(None, 4, 0, 0x10, 10, 4, 0x20, 4, 5, 0, k, 0, None, 0, None) == _abi.EINVAL  # src NULL
    assert f(0x10, 4, 0, None, 10, 4, 0x20, 4, 5, 0, k, 0, None, 0, None) == _abi.EINVAL  # index NULL
    assert f(0x10, 4, 0, 0x10, 10, 4, None, 4, 5, 0, k, 0, None, 0, None) == _abi.EINVAL  # dst NULL
    assert f(0x10, 2, 0, 0x10, 10, 4, 0x20, 4, 5, 0, k, 0, None, 0, None) == _abi.EINVAL  # ld_src < N
    assert f(0x10, 4, 0, 0x10, 10, 4, 0x20, 4, 0, 0, k, 0, None, 0, None) == _abi.EINVAL  # dst_len 0
    assert f(0x10, 4, 0, 0x10, -1, 4, 0x20, 4, 5, 0, k, 0, None, 0, None) == _abi.EINVAL  # E < 0
    assert f(0x10, 4, 0, 0x10, 10, 4, 0x20, 4, 5, 99, k, 0, None, 0, None) == _abi.EINVAL  # variant
    kk = _abi.Knobs.make(write_policy=2)
    assert f(0x10, 4, 0, 0x10, 10, 4, 0x20, 4, 5, _abi.SEG_SR, byref(kk), 0, None, 0, None) == _abi.EWORKSPACE
    kk = _abi.Knobs.make(dtile_size=3)
    assert f(0x10, 4, 0, 0x10, 10, 4, 0x20, 4, 5, _abi.SEG_SORTED_SR, byref(kk), 0, None, 0, None) == _abi.EUNSUPPORTED


def test_backward_validation(lib):
    d = _desc()
    assert lib.segscan_csr_transpose(byref(d), None, 0x10, 0x10, None, 0x10, 1 << 20, None) == _abi.EINVAL
    assert lib.segscan_csr_transpose(byref(d), 0x10, 0x10, 0x10, None, None, 0, None) == _abi.EWORKSPACE
    assert lib.segscan_sddmm_csr(byref(d), 8, None, 8, 0x10, 8, 0, 0x10, None) == _abi.EINVAL
    assert lib.segscan_segment_gather(None, 8, 0x10, 10, 8, None, 0x10, 8, None) == _abi.EINVAL
    assert lib.segscan_segment_counts(None, 10, 5, 0x10, None) == _abi.EINVAL
    assert lib.segscan_check_csr(byref(d), None, None) == _abi.EINVAL


def test_fused_backward_validation_and_plan(lib):
    d = _desc(nrow=10, ncol=12, nnz=20)
    f = lib.segscan_spmm_backward_csr

    def call(perm=0x10, N=8, dC=0x100, ld_dc=8, dc_t=_abi.F32, B=0x200, ld_b=8, b_t=_abi.F32, dB=0x300, ld_db=8,
             dvals=0x400, knobs=None, flags=0, ws=0x500, ws_bytes=1 << 30):
        return f(byref(d), perm, None, N, dC, ld_dc, dc_t, B, ld_b, b_t, dB, ld_db, dvals, knobs, flags, ws, ws_bytes,
                 None)

    assert call(perm=None) == _abi.EINVAL
    assert call(dC=None) == _abi.EINVAL
    assert call(ld_dc=4) == _abi.EINVAL
    assert call(ld_db=4) == _abi.EINVAL
    assert call(flags=4) == _abi.EINVAL
    assert call(ws=None, ws_bytes=0) == _abi.EWORKSPACE
    assert call(dc_t=9) == _abi.EINVAL
    # a 16-bit dC needs a B of the same type; an fp32 dC takes any B
    assert call(dc_t=_abi.BF16, b_t=_abi.F32) == _abi.EUNSUPPORTED
    assert call(dc_t=_abi.BF16, b_t=_abi.F16) == _abi.EUNSUPPORTED
    # wider than one group of 64 lanes x 8 columns
    assert call(N=600, ld_dc=600, ld_b=600, ld_db=600) == _abi.EUNSUPPORTED
    assert "coarsen_factor" in lib.segscan_last_error().decode()
    assert call(knobs=byref(_abi.Knobs.make(thread_per_block=512))) == _abi.EUNSUPPORTED
    # workspace: one carry row index and one N-wide carry per chunk
    n = c_size_t(0)
    big = _desc(nrow=262144, ncol=262144, nnz=4194304)
    assert lib.segscan_spmm_backward_workspace_size(byref(big), 64, _abi.F32, None, byref(n)) == _abi.OK
    chunks = 4194304 // (16 * 16)  # 16 lanes x 4 columns cover N = 64; thread_nz 16 at this nnz
    assert n.value == chunks * 4 + chunks * 64 * 4
    kk = _abi.Knobs.make(thread_nz=2)
    assert lib.segscan_spmm_backward_workspace_size(byref(big), 64, _abi.F32, byref(kk), byref(n)) == _abi.OK
    assert n.value == (4194304 // 32) * 4 + (4194304 // 32) * 64 * 4


def test_fused_backward_rowmap_validation(lib):
    """The row-map entry point needs the map and otherwise validates as the plain one."""
    d = _desc(nrow=10, ncol=12, nnz=20)
    f = lib.segscan_spmm_backward_csr_rowmap

    def call(perm=0x10, N=8, dC=0x100, dc_map=0x180, dc_t=_abi.F32, B=0x200, dB=0x300, ws=0x500):
        return f(byref(d), perm, None, N, dC, 8, dc_t, dc_map, B, 8, _abi.F32, dB, 8, 0x400, None, 0, ws, 1 << 30,
                 None)

    assert call(dc_map=None) == _abi.EINVAL
    assert "dc_map" in lib.segscan_last_error().decode()
    assert call(perm=None) == _abi.EINVAL
    assert call(dC=None) == _abi.EINVAL
    assert call(dc_t=_abi.BF16) == _abi.EUNSUPPORTED  # bf16 dC with an fp32 B
    assert call(ws=None) == _abi.EWORKSPACE
    # nothing to gather (N = 0): no map needed, dvals zero-filled on the device
    e = _desc(nrow=10, ncol=12, nnz=0)
    assert f(byref(e), None, None, 8, None, 8, _abi.F32, None, None, 8, _abi.F32, None, 8, None, None, 0, None, 0,
             None) == _abi.OK


def _describe(lib, nrow, nnz, N, variant, dtype=_abi.F32, max_row=0, **knobs):
    info = _abi.LaunchInfo()
    d = _desc(nrow=nrow, ncol=nrow, nnz=nnz, max_row=max_row)
    kp = byref(_abi.Knobs.make(**knobs)) if knobs else None
    st = lib.segscan_spmm_describe(byref(d), N, dtype, variant, kp, byref(info))
    assert st == _abi.OK, lib.segscan_last_error().decode()
    return info


def test_describe_plain_ebpr_geometry(lib):
    # spmm.cuh:172-187 at N=64: CF 4, 8 panels x 32 lanes, gridDim.y 2
    info = _describe(lib, 262144, 4194304, 64, _abi.SPMM_PLAIN_EBPR)
    assert (info.group_size, info.tile_size, info.coarsen_factor) == (32, 32, 4)
    assert info.thread_per_block == 256 and info.grid_y == 2
    assert info.write_policy == _abi.WRITE_ATOMIC
    assert info.nchunks == 4194304 // 32  # tiles of 32 nnz, one partition entry each


def test_describe_cg_defaults_and_knobs(lib):
    info = _describe(lib, 262144, 4194304, 64, _abi.SPMM_EBSR)  # (4,256,4,1,1,16)
    assert (info.group_size, info.coarsen_factor, info.thread_nz) == (16, 4, 1)
    assert info.grid_y == 1
    info = _describe(lib, 262144, 4194304, 64, _abi.SPMM_EBSR, group_factor=4, thread_nz=16, write_policy=2)
    assert info.chunk == 256 and info.nchunks == 4194304 // 256
    assert info.workspace_bytes >= info.nchunks * (4 + 64 * 4)  # owner carries: row + N values per chunk
    info = _describe(lib, 262144, 4194304, 64, _abi.SPMM_RBPR)  # (3,256,3,1,2): G == T -> stores
    assert info.write_policy == _abi.WRITE_OWNER and info.workspace_bytes == 0
    info = _describe(lib, 262144, 4194304, 64, _abi.SPMM_RBPR, group_factor=2, tile_factor=3)
    assert info.write_policy == _abi.WRITE_ATOMIC
    # bf16 storage doubles the default columns per lane (128-bit loads)
    info = _describe(lib, 262144, 4194304, 64, _abi.SPMM_PLAIN_EBPR, dtype=_abi.BF16)
    assert info.coarsen_factor == 8


def test_describe_narrows_oversized_slices(lib):
    # tile 64 x one-column panels would need 64 x 64 lanes: the slice shrinks to 16 columns
    info = _describe(lib, 1000, 5000, 64, _abi.SPMM_EBPR, group_factor=6, tile_factor=6, coarsen_factor=1)
    assert info.thread_per_block == 1024 and info.grid_y == 4


def test_describe_auto_is_concrete(lib):
    info = _describe(lib, 262144, 4194304, 64, _abi.SPMM_AUTO)
    assert info.variant in range(4)
    assert info.thread_per_block % 64 == 0 and info.thread_per_block <= 1024


@pytest.mark.parametrize("nrow,nnz,max_row,N,variant,group,tnz", [
    # even rows (longest <= 4 x mean + 32): one group per row
    (16384, 65536, 14, 4, _abi.SPMM_RBPR, 4, None),      # mean 4 -> groups of 4 lanes
    (16384, 262144, 40, 4, _abi.SPMM_RBPR, 8, None),     # mean 16 -> 8
    (16384, 1048576, 96, 8, _abi.SPMM_RBPR, 16, None),   # mean 64 -> 16
    (16384, 524288, 60, 16, _abi.SPMM_RBPR, 16, None),   # mean 32, N 16 -> still a group per row
    (16384, 65536, 14, 16, _abi.SPMM_RBSR, None, None),
    (16384, 65536, 14, 64, _abi.SPMM_RBSR, None, None),
    (16384, 1048576, 90, 128, _abi.SPMM_RBSR, None, None),
    (16384, 1048576, 90, 64, _abi.SPMM_RBSR, None, None),
    # skewed or unknown: element-balanced, 4-byte slices per lane (one fp32 column) in groups of
    # up to 64 lanes, thread_nz = 2^floor(log2(1.125 nnz / 2^18)) within [64 / group, 32]
    (16384, 65536, 0, 64, _abi.SPMM_EBSR, 64, 1),
    (262144, 4194295, 20000, 64, _abi.SPMM_EBSR, 64, 16),   # just under 2^22: still 16
    (262144, 2097013, 20000, 64, _abi.SPMM_EBSR, 64, 8),
    (65536, 588341, 65536, 64, _abi.SPMM_EBSR, 64, 2),
    (262144, 4194304, 20000, 64, _abi.SPMM_EBSR, 64, 16),
    (262144, 4194304, 20000, 128, _abi.SPMM_EBSR, 64, 16),  # two columns per lane
    (65536, 1048576, 20000, 16, _abi.SPMM_EBSR, 16, 4),
    (65536, 1048576, 20000, 4, _abi.SPMM_EBSR, 4, 16),
    (262144, 2**28, 20000, 64, _abi.SPMM_EBSR, 64, 32),
])
def test_describe_auto_follows_row_hint(lib, nrow, nnz, max_row, N, variant, group, tnz):
    info = _describe(lib, nrow, nnz, N, _abi.SPMM_AUTO, max_row=max_row)
    assert info.variant == variant
    if group is not None:
        assert info.group_size == group
    if tnz is not None:
        assert info.thread_nz == tnz
    if variant == _abi.SPMM_EBSR:  # up to 32 waves per CU on 256 CUs, owner stores
        assert info.grid_x <= 2048 and info.write_policy == _abi.WRITE_OWNER and info.chunk >= 64


def test_describe_auto_16bit_takes_two_columns_per_lane(lib):
    for N, group in ((64, 32), (128, 64)):
        info = _describe(lib, 262144, 4194304, N, _abi.SPMM_AUTO, dtype=_abi.BF16, max_row=20000)
        assert (info.variant, info.group_size, info.coarsen_factor) == (_abi.SPMM_EBSR, group, 2)
    info = _describe(lib, 262144, 4194304, 512, _abi.SPMM_AUTO, dtype=_abi.F32, max_row=20000)
    assert (info.group_size, info.coarsen_factor) == (64, 8)  # widened only when N needs more than 64 lanes


def test_describe_auto_even_rows_take_8_byte_slices(lib):
    """Even rows: row-balanced sequential reduction over 64-column slices with at most 4
    columns per lane, for fp32 and 16-bit B alike (16 lanes per row from N = 64 on), narrower
    until a row has at least 8 lanes."""
    for dtype in (_abi.F32, _abi.BF16, _abi.F16):
        for N, cf in ((64, 4), (128, 4), (32, 4), (20, 2), (16, 2), (34, 2), (33, 1), (24, 2)):
            info = _describe(lib, 120000, 2520000, N, _abi.SPMM_AUTO, dtype=dtype, max_row=24)
            assert info.variant == _abi.SPMM_RBSR, (dtype, N)
            assert info.coarsen_factor == cf, (dtype, N, info.coarsen_factor)


def test_workspace_size_query(lib):
    d = _desc(nrow=1000, ncol=1000, nnz=5000)
    n = c_size_t(0)
    assert lib.segscan_spmm_workspace_size(byref(d), 16, _abi.F32, _abi.SPMM_PLAIN_RBSR, None, byref(n)) == _abi.OK
    assert n.value == 0
    assert lib.segscan_spmm_workspace_size(byref(d), 16, _abi.F32, _abi.SPMM_EBPR, None, byref(n)) == _abi.OK
    assert n.value > 0
    assert lib.segscan_csr_transpose_workspace_size(byref(d), byref(n)) == _abi.OK
    assert n.value >= 1000 * 4


def test_batch_plan_validation(lib):
    """segscan_batch_plan_create rejects bad arguments on the host, before any HIP call."""
    offs = (ctypes.c_int64 * 3)(0, 4, 10)
    plan = ctypes.c_void_p()

    def create(desc, offsets=offs, n_graphs=2, N=16, dtype=_abi.F32, knobs=None):
        kp = None if knobs is None else byref(_abi.Knobs.make(**knobs))
        return lib.segscan_batch_plan_create(None if desc is None else byref(desc), offsets, n_graphs, N, dtype, kp,
                                             None, byref(plan))

    sq = _desc(nrow=10, ncol=10, nnz=20)
    assert create(None) == _abi.EINVAL
    assert create(_desc(nrow=10, ncol=12, nnz=20)) == _abi.EINVAL          # not square
    assert create(sq, offsets=None) == _abi.EINVAL
    assert create(sq, offsets=(ctypes.c_int64 * 3)(0, 4, 9)) == _abi.EINVAL  # does not end at M
    assert create(sq, offsets=(ctypes.c_int64 * 3)(0, 6, 4)) == _abi.EINVAL  # decreasing (and ends wrong)
    assert create(sq, offsets=(ctypes.c_int64 * 3)(1, 4, 10)) == _abi.EINVAL
    assert create(sq, n_graphs=0) == _abi.EINVAL
    assert create(sq, N=0) == _abi.EUNSUPPORTED
    assert create(sq, N=513) == _abi.EUNSUPPORTED
    assert create(sq, N=300) == _abi.EUNSUPPORTED                          # fp32: at most 64 lanes x 4 columns
    assert create(sq, dtype=7) == _abi.EINVAL
    assert create(sq, N=64, knobs=dict(group_factor=3, coarsen_factor=1)) == _abi.EUNSUPPORTED
    assert not plan.value
    # a NULL plan is refused by every call but destroy
    info = _abi.BatchInfo()
    assert lib.segscan_batch_plan_describe(None, byref(info)) == _abi.EINVAL
    assert lib.segscan_batch_plan_run(None, 0x1000, 16, None, 0, None, 0, 0x2000, 16, None, 0, None, None) == _abi.EINVAL
    assert lib.segscan_batch_plan_set_values(None, None, None) == _abi.EINVAL
    assert lib.segscan_batch_plan_destroy(None) == _abi.OK


def test_values_order_validation(lib):
    """vals_map needs values; the transpose and the backward's At refuse one; the backward
    knows the At-order dvals flag; the inverse permutation checks its arguments."""
    d = _desc(nrow=10, ncol=10, nnz=20)
    d.vals_map = 0x3000
    assert _spmm(lib, d) == _abi.EINVAL                      # a map without values
    d.vals = 0x4000
    assert _spmm(lib, d, variant=_abi.SPMM_EBSR, ws=None) in (_abi.EWORKSPACE, _abi.OK, _abi.EHIP)
    ws = ctypes.c_size_t()
    assert lib.segscan_csr_transpose_workspace_size(byref(d), byref(ws)) == _abi.OK
    assert lib.segscan_csr_transpose(byref(d), 0x5000, 0x6000, 0x7000, None, 0x8000, ws.value, None) == \
        _abi.EUNSUPPORTED
    st = lib.segscan_spmm_backward_csr(byref(d), 0x9000, None, 8, 0xa000, 8, _abi.F32, 0xb000, 8, _abi.F32, None, 0,
                                       0xc000, None, 0, None, 0, None)
    assert st == _abi.EINVAL                                  # At.vals_map must be NULL
    d.vals_map = None
    st = lib.segscan_spmm_backward_csr(byref(d), 0x9000, None, 8, 0xa000, 8, _abi.F32, 0xb000, 8, _abi.F32, None, 0,
                                       0xc000, None, 4, None, 0, None)
    assert st == _abi.EINVAL                                  # unknown flag
    assert lib.segscan_invert_permutation(None, -1, None, None) == _abi.EINVAL
    assert lib.segscan_invert_permutation(None, 0, None, None) == _abi.OK
    assert lib.segscan_invert_permutation(None, 5, None, None) == _abi.EINVAL


def test_gather_probe_validation(lib):
    n = ctypes.c_int64()
    assert lib.segscan_probe_gather_out_len(64, 100, 16, byref(n)) == _abi.EINVAL   # 128 or 256 bytes only
    assert lib.segscan_probe_gather_out_len(256, 100, 0, byref(n)) == _abi.EINVAL
    assert lib.segscan_probe_gather_out_len(256, 1000, 16, byref(n)) == _abi.OK
    assert n.value == 1024  # 63 groups x 16 lanes, rounded up to whole workgroups of 256
    assert lib.segscan_probe_gather(0x1000, 10, 256, 0x2000, 1000, 16, 0x3000, 100, None) == _abi.EINVAL  # out short
    assert lib.segscan_probe_gather(0x1008, 10, 256, 0x2000, 1000, 16, 0x3000, 1024, None) == _abi.EINVAL  # misaligned


def test_permute_plan_validation(lib):
    """segscan_permute_plan_create / run reject bad arguments on the host, before any HIP call
    (whether src is a permutation is checked after reading it back: GPU tests)."""
    h = ctypes.c_void_p()
    assert lib.segscan_permute_plan_create(0x1000, 0, None, byref(h)) == _abi.EINVAL
    assert lib.segscan_permute_plan_create(0x1000, -5, None, byref(h)) == _abi.EINVAL
    assert lib.segscan_permute_plan_create(0x1000, 1 << 31, None, byref(h)) == _abi.EINVAL
    assert lib.segscan_permute_plan_create(None, 16, None, byref(h)) == _abi.EINVAL
    assert lib.segscan_permute_plan_create(0x1000, 16, None, None) == _abi.EINVAL
    assert lib.segscan_last_error().decode() != ""
    assert lib.segscan_permute_run(None, 0x1000, 0x2000, None) == _abi.EINVAL
    assert lib.segscan_permute_plan_destroy(None) == _abi.OK


def test_spmm_permuted_host_validation(lib):
    """segscan_spmm_csr_permuted refuses a missing A or plan, and an A that carries its own
    values or value map, on the host before touching the plan or the device (the plan's own
    checks and the launch routes: GPU tests)."""
    d = _abi.CsrDesc()
    d.nrow, d.ncol, d.nnz = 4, 4, 4
    d.rowptr, d.colind = 0x1000, 0x2000

    def call(desc, plan):
        return lib.segscan_spmm_csr_permuted(desc, plan, 0x3000, 8, 0x4000, 8, _abi.F32, 0x5000, 8,
                                            _abi.SPMM_AUTO, None, 0, None, 0, None)

    assert call(None, 0x6000) == _abi.EINVAL
    assert call(byref(d), None) == _abi.EINVAL and "plan" in lib.segscan_last_error().decode()
    d.vals = 0x7000
    assert call(byref(d), 0x6000) == _abi.EINVAL and "must be NULL" in lib.segscan_last_error().decode()
    d.vals, d.vals_map = None, 0x7000
    assert call(byref(d), 0x6000) == _abi.EINVAL and "must be NULL" in lib.segscan_last_error().decode()
