"""ctypes binding of the C ABI declared in include/segscan.h.

Loading fails loudly (ImportError) when libsegscan_hip.so has not been built:
there is no CPU fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_int32, c_int64, c_size_t, c_uint32, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libsegscan_hip.so"

# segscan_status_t
OK, EINVAL, EUNSUPPORTED, EHIP, EWORKSPACE = 0, 1, 2, 3, 4
# segscan_dtype_t
F32, BF16, F16 = 0, 1, 2
# segscan_spmm_variant_t
SPMM_EBPR, SPMM_RBPR, SPMM_EBSR, SPMM_RBSR = 0, 1, 2, 3
SPMM_PLAIN_EBPR, SPMM_PLAIN_RBPR, SPMM_PLAIN_EBSR, SPMM_PLAIN_RBSR = 4, 5, 6, 7
SPMM_AUTO = 8
# segscan_seg_variant_t
SEG_PR, SEG_PR_NEW, SEG_SR, SEG_NOSHMEM_SR, SEG_LESSATOM_SR, SEG_SORTED_SR, SEG_AUTO = range(7)
# segscan_write_policy_t
WRITE_DEFAULT, WRITE_ATOMIC, WRITE_OWNER, WRITE_LESSATOM = 0, 1, 2, 3
ACCUMULATE = 1
DVALS_AT_ORDER = 2
# segscan_mtx_read flags
MTX_KEEP_VALUES, MTX_REFERENCE_QUIRKS = 1, 2

SPMM_VARIANTS = {
    "ebpr": SPMM_EBPR, "rbpr": SPMM_RBPR, "ebsr": SPMM_EBSR, "rbsr": SPMM_RBSR,
    "plain_ebpr": SPMM_PLAIN_EBPR, "plain_rbpr": SPMM_PLAIN_RBPR,
    "plain_ebsr": SPMM_PLAIN_EBSR, "plain_rbsr": SPMM_PLAIN_RBSR, "auto": SPMM_AUTO,
}
SEG_VARIANTS = {
    "segment_coo": SEG_PR, "segment_coo_new": SEG_PR_NEW, "segment_coo_sr": SEG_SR,
    "segment_coo_noshmem_sr": SEG_NOSHMEM_SR, "segment_coo_noshmem_lessatom_sr": SEG_LESSATOM_SR,
    "segment_coo_sorted_sr": SEG_SORTED_SR, "auto": SEG_AUTO,
}


class CsrDesc(ctypes.Structure):
    _fields_ = [
        ("nrow", c_int64), ("ncol", c_int64), ("nnz", c_int64),
        ("rowptr", c_void_p), ("colind", c_void_p), ("vals", c_void_p),
        ("max_row_nnz", c_int64), ("vals_map", c_void_p),
    ]


class HostCsrDesc(ctypes.Structure):
    _fields_ = [
        ("nrow", c_int64), ("ncol", c_int64), ("nnz", c_int64),
        ("rowptr", POINTER(c_int32)), ("colind", POINTER(c_int32)), ("vals", POINTER(ctypes.c_float)),
    ]


KNOB_FIELDS = (
    "group_factor", "tile_factor", "thread_per_block", "block_numer", "block_denom",
    "coarsen_factor", "thread_nz", "dtile_size", "dtile_num", "block_dim_y", "write_policy",
)


class Knobs(ctypes.Structure):
    _fields_ = [(name, c_int32) for name in KNOB_FIELDS]

    @classmethod
    def make(cls, **kw) -> "Knobs":
        unknown = set(kw) - set(KNOB_FIELDS)
        if unknown:
            raise TypeError(f"unknown knobs: {sorted(unknown)}")
        k = cls()
        for name, value in kw.items():
            if value is None:
                continue
            if name in ("group_factor", "tile_factor") and value == 0:
                value = -1  # explicit size 1 (0 means "default" in the C struct)
            setattr(k, name, int(value))
        return k

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name in KNOB_FIELDS}


class LaunchInfo(ctypes.Structure):
    _fields_ = [
        ("variant", c_int32), ("group_size", c_int32), ("tile_size", c_int32),
        ("thread_per_block", c_int32), ("coarsen_factor", c_int32), ("thread_nz", c_int32),
        ("write_policy", c_int32), ("chunk", c_int32), ("grid_x", c_int64), ("grid_y", c_int64),
        ("nchunks", c_int64), ("workspace_bytes", c_size_t),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchInfo(ctypes.Structure):
    _fields_ = [
        ("nitems", c_int64), ("grid", c_int64), ("group_size", c_int32), ("coarsen_factor", c_int32),
        ("thread_per_block", c_int32), ("item_target", c_int64), ("plan_bytes", c_size_t),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class SegscanError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{message} (status {status})")
        self.status = status


_SIGNATURES = {
    "segscan_abi_version": (c_int32, []),
    "segscan_status_string": (c_char_p, [c_int32]),
    "segscan_last_error": (c_char_p, []),
    "segscan_spmm_describe": (c_int32, [POINTER(CsrDesc), c_int64, c_int32, c_int32, POINTER(Knobs), POINTER(LaunchInfo)]),
    "segscan_spmm_workspace_size": (c_int32, [POINTER(CsrDesc), c_int64, c_int32, c_int32, POINTER(Knobs), POINTER(c_size_t)]),
    "segscan_spmm_csr": (c_int32, [POINTER(CsrDesc), c_int64, c_void_p, c_int64, c_int32, c_void_p, c_int64, c_int32,
                                   POINTER(Knobs), c_uint32, c_void_p, c_size_t, c_void_p]),
    "segscan_batch_plan_create": (c_int32, [POINTER(CsrDesc), POINTER(c_int64), c_int64, c_int64, c_int32,
                                            POINTER(Knobs), c_void_p, POINTER(c_void_p)]),
    "segscan_batch_plan_set_values": (c_int32, [c_void_p, c_void_p, c_void_p]),
    "segscan_batch_plan_describe": (c_int32, [c_void_p, POINTER(BatchInfo)]),
    "segscan_batch_plan_run": (c_int32, [c_void_p, c_void_p, c_int64, c_void_p, c_int64, c_void_p, c_int64, c_void_p,
                                         c_int64, c_void_p, c_int64, c_void_p, c_void_p]),
    "segscan_batch_plan_destroy": (c_int32, [c_void_p]),
    "segscan_tune_prefix": (c_char_p, []),
    "segscan_tune_table_create": (c_int32, [POINTER(c_void_p)]),
    "segscan_tune_table_destroy": (c_int32, [c_void_p]),
    "segscan_tune_table_load": (c_int32, [c_void_p, c_char_p, POINTER(c_int64), POINTER(c_int64)]),
    "segscan_tune_table_save": (c_int32, [c_void_p, c_char_p]),
    "segscan_tune_table_lookup": (c_int32, [c_void_p, POINTER(CsrDesc), c_int64, c_int32, POINTER(c_int32),
                                            POINTER(Knobs), POINTER(c_int32)]),
    "segscan_autotune_workspace_size": (c_int32, [POINTER(CsrDesc), c_int64, c_int32, POINTER(c_size_t)]),
    "segscan_autotune_spmm": (c_int32, [c_void_p, POINTER(CsrDesc), c_int64, c_void_p, c_int64, c_int32, c_void_p,
                                        c_int64, c_int32, c_void_p, c_size_t, c_void_p, POINTER(c_int32),
                                        POINTER(Knobs), POINTER(ctypes.c_float)]),
    "segscan_probe_gather_out_len": (c_int32, [c_int64, c_int64, c_int64, POINTER(c_int64)]),
    "segscan_probe_gather": (c_int32, [c_void_p, c_int64, c_int64, c_void_p, c_int64, c_int64, c_void_p, c_int64,
                                       c_void_p]),
    "segscan_invert_permutation": (c_int32, [c_void_p, c_int64, c_void_p, c_void_p]),
    "segscan_permute_plan_create": (c_int32, [c_void_p, c_int64, c_void_p, POINTER(c_void_p)]),
    "segscan_permute_run": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "segscan_spmm_csr_permuted": (c_int32, [POINTER(CsrDesc), c_void_p, c_void_p, c_int64, c_void_p, c_int64, c_int32,
                                            c_void_p, c_int64, c_int32, POINTER(Knobs), c_uint32, c_void_p, c_size_t,
                                            c_void_p]),
    "segscan_permute_plan_destroy": (c_int32, [c_void_p]),
    "segscan_spmm_csr_noalloc": (c_int32, [POINTER(CsrDesc), c_int64, c_void_p, c_int64, c_int32, c_void_p, c_int64,
                                           c_int32, POINTER(Knobs), c_uint32, c_void_p]),
    "segscan_segment_coo_describe": (c_int32, [c_int64, c_int64, c_int64, c_int32, c_int32, POINTER(Knobs), POINTER(LaunchInfo)]),
    "segscan_segment_coo_workspace_size": (c_int32, [c_int64, c_int64, c_int64, c_int32, c_int32, POINTER(Knobs), POINTER(c_size_t)]),
    "segscan_segment_coo": (c_int32, [c_void_p, c_int64, c_int32, c_void_p, c_int64, c_int64, c_void_p, c_int64, c_int64,
                                      c_int32, POINTER(Knobs), c_uint32, c_void_p, c_size_t, c_void_p]),
    "segscan_csr_transpose_workspace_size": (c_int32, [POINTER(CsrDesc), POINTER(c_size_t)]),
    "segscan_csr_transpose": (c_int32, [POINTER(CsrDesc), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "segscan_sddmm_csr": (c_int32, [POINTER(CsrDesc), c_int64, c_void_p, c_int64, c_void_p, c_int64, c_int32, c_void_p, c_void_p]),
    "segscan_spmm_backward_workspace_size": (c_int32, [POINTER(CsrDesc), c_int64, c_int32, POINTER(Knobs), POINTER(c_size_t)]),
    "segscan_spmm_backward_csr": (c_int32, [POINTER(CsrDesc), c_void_p, c_void_p, c_int64, c_void_p, c_int64, c_int32,
                                            c_void_p, c_int64, c_int32, c_void_p, c_int64, c_void_p, POINTER(Knobs),
                                            c_uint32, c_void_p, c_size_t, c_void_p]),
    "segscan_spmm_backward_csr_rowmap": (c_int32, [POINTER(CsrDesc), c_void_p, c_void_p, c_int64, c_void_p, c_int64,
                                                   c_int32, c_void_p, c_void_p, c_int64, c_int32, c_void_p, c_int64,
                                                   c_void_p, POINTER(Knobs), c_uint32, c_void_p, c_size_t, c_void_p]),
    "segscan_segment_gather": (c_int32, [c_void_p, c_int64, c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_int64, c_void_p]),
    "segscan_segment_counts": (c_int32, [c_void_p, c_int64, c_int64, c_void_p, c_void_p]),
    "segscan_check_csr": (c_int32, [POINTER(CsrDesc), c_void_p, c_void_p]),
    "segscan_check_index": (c_int32, [c_void_p, c_int64, c_int64, c_void_p, c_void_p]),
    "segscan_set_timing_events": (None, [c_void_p, c_void_p]),
    "segscan_mtx_read": (c_int32, [c_char_p, c_uint32, c_void_p]),
    "segscan_host_csr_free": (None, [c_void_p]),
    "segscan_mtx_last_error": (c_char_p, []),
}

EXPORTED = tuple(_SIGNATURES)

_lib = None


def load() -> ctypes.CDLL:
    """Loads the HIP library once; raises ImportError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("SEGSCAN_LIB", LIB_PATH))
    if not path.exists():
        raise ImportError(
            f"{path} is missing: build the HIP library first (python -m segscan_amd.build); "
            "segscan_amd has no CPU fallback")
    lib = ctypes.CDLL(str(path))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int, what: str = "segscan") -> None:
    if status != OK:
        lib = load()
        msg = lib.segscan_last_error().decode() or lib.segscan_status_string(status).decode()
        raise SegscanError(status, f"{what}: {msg}")


def knobs_ptr(knobs):
    if knobs is None:
        return None
    if isinstance(knobs, Knobs):
        return ctypes.byref(knobs)
    if isinstance(knobs, dict):
        return ctypes.byref(Knobs.make(**knobs))
    raise TypeError("knobs must be a Knobs, a dict or None")
