"""Matrix Market input through the library's host reader (segscan_mtx_read).

``read_mtx(path)`` returns the CSR the reference's read_mtx_file builds
(test-cg/include/dataloader/dataloader.hpp:22-107) with values set to one
(:154-155); ``keep_values=True`` keeps the file's values instead.
``reference_quirks=True`` also reproduces its two deviations: only "symmetric"
files are mirrored, and the CSR is cut to the declared entry count (:95).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _abi
from .synth import HostCsr


class MtxError(ValueError):
    pass


def read_mtx(path: str | os.PathLike, keep_values: bool = False, reference_quirks: bool = False) -> HostCsr:
    lib = _abi.load()
    out = _abi.HostCsrDesc()
    flags = (_abi.MTX_KEEP_VALUES if keep_values else 0) | (_abi.MTX_REFERENCE_QUIRKS if reference_quirks else 0)
    st = lib.segscan_mtx_read(os.fsencode(path), flags, ctypes.byref(out))
    if st != _abi.OK:
        raise MtxError(f"{os.fspath(path)}: {lib.segscan_mtx_last_error().decode()}")
    try:
        nrow, nnz = int(out.nrow), int(out.nnz)
        rowptr = np.ctypeslib.as_array(out.rowptr, shape=(nrow + 1,)).copy()
        colind = np.ctypeslib.as_array(out.colind, shape=(max(nnz, 1),))[:nnz].copy()
        vals = np.ctypeslib.as_array(out.vals, shape=(max(nnz, 1),))[:nnz].copy()
    finally:
        lib.segscan_host_csr_free(ctypes.byref(out))
    return HostCsr(nrow, int(out.ncol), rowptr, colind, vals)
