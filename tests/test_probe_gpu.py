"""The gather probe reads the rows it is given: its per-lane checksums equal a numpy
restatement (for each group's chunk of indices, lane l sums the xor of its four 32-bit
words of each row, modulo 2^32, and stores the low 16 bits as a float)."""
import ctypes

import numpy as np
import pytest
import torch

from segscan_amd import _abi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("row_bytes,chunk", [(256, 16), (256, 1024), (128, 64)])
def test_probe_checksums(gpu, row_bytes, chunk):
    lib = _abi.load()
    rng = np.random.default_rng(row_bytes + chunk)
    nrows, n = 5000, 30011
    table = rng.integers(0, 2**32, size=(nrows, row_bytes // 4), dtype=np.uint64).astype(np.uint32)
    idx = rng.integers(0, nrows, size=n).astype(np.int32)
    out_len = ctypes.c_int64()
    _abi.check(lib.segscan_probe_gather_out_len(row_bytes, n, chunk, ctypes.byref(out_len)), "len")
    t = torch.from_numpy(table.view(np.int32)).to(gpu)
    i = torch.from_numpy(idx).to(gpu)
    out = torch.full((out_len.value,), -1.0, device=gpu)
    _abi.check(lib.segscan_probe_gather(ctypes.c_void_p(t.data_ptr()), nrows, row_bytes, ctypes.c_void_p(i.data_ptr()),
                                        n, chunk, ctypes.c_void_p(out.data_ptr()), out_len.value,
                                        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "probe")
    got = out.cpu().numpy()
    G = row_bytes // 16
    words = table.reshape(nrows, G, 4)
    x = (words[:, :, 0] ^ words[:, :, 1] ^ words[:, :, 2] ^ words[:, :, 3]).astype(np.uint64)  # [rows, G]
    groups = -(-n // chunk)
    want = np.zeros(out_len.value, np.float32)
    for gi in range(groups):
        rows = idx[gi * chunk:(gi + 1) * chunk]
        acc = (x[rows].sum(axis=0) & 0xFFFFFFFF) & 0xFFFF
        want[gi * G:(gi + 1) * G] = acc.astype(np.float32)
    # lanes past the last group sum nothing
    assert np.array_equal(got[:groups * G], want[:groups * G])
    assert np.all(got[groups * G:] == 0)
