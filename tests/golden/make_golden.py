"""Writes the golden fixtures of tests/golden/ from computations independent of
the oracle (dense float64 numpy products, np.add.at, bisect, numpy's stable
argsort). The reference ships no expected outputs for this path (its checks
compare GPU and CPU runs at run time), so these fixtures fix small cases by
textbook routines instead. Run: python tests/golden/make_golden.py
"""
import bisect
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def dense_of(M, K, rowptr, colind, vals):
    D = np.zeros((M, K), np.float64)
    for i in range(M):
        for p in range(rowptr[i], rowptr[i + 1]):
            D[i, colind[p]] += vals[p]
    return D


def main():
    rng = np.random.default_rng(20240601)

    # 1. hand-written CSR: empty first row, empty middle rows, trailing empty row
    rowptr = np.array([0, 0, 2, 5, 5, 5, 7, 8, 8], np.int32)
    colind = np.array([1, 4, 0, 2, 3, 0, 4, 2], np.int32)
    vals = np.array([0.5, -1.25, 2.0, 0.75, -0.5, 1.5, -2.0, 0.25], np.float32)
    B = np.array([[1, -1, 0.5], [2, 0, -2], [-1, 1, 1], [0.5, 0.5, 0.5], [3, -3, 1]], np.float32)
    C = dense_of(8, 5, rowptr, colind, vals) @ B.astype(np.float64)
    np.savez(HERE / "spmm_tiny.npz", rowptr=rowptr, colind=colind, vals=vals, B=B, C=C)

    # 2. random integer-valued CSR (exact in any summation order)
    M, K, N = 64, 48, 7
    deg = rng.integers(0, 12, size=M)
    deg[[3, 17, 40]] = 0
    rp = np.concatenate(([0], np.cumsum(deg))).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(K, size=d, replace=False)) for d in deg]).astype(np.int32)
    va = rng.integers(-2, 3, size=ci.size).astype(np.float32)
    Bi = rng.integers(-3, 4, size=(K, N)).astype(np.float32)
    Ci = dense_of(M, K, rp, ci, va) @ Bi.astype(np.float64)
    np.savez(HERE / "spmm_int.npz", rowptr=rp, colind=ci, vals=va, B=Bi, C=Ci)

    # 3. segment_coo with empty segments at the start, middle and end
    runs = np.array([0, 3, 1, 0, 0, 7, 2, 0, 5, 0], np.int64)
    index = np.repeat(np.arange(runs.size), runs).astype(np.int32)
    src = rng.integers(-5, 6, size=(index.size, 5)).astype(np.float32)
    dst = np.zeros((runs.size, 5), np.float64)
    np.add.at(dst, index, src.astype(np.float64))
    np.savez(HERE / "segment_int.npz", index=index, src=src, dst=dst)

    # 4. row search: rowptr with runs of empty rows, every query position
    rp4 = np.array([0, 0, 0, 3, 3, 4, 9, 9, 9, 10], np.int32)
    nnz = int(rp4[-1])
    queries = np.arange(0, nnz + 3, dtype=np.int64)
    M4 = rp4.size - 1
    rows = np.array([min(bisect.bisect_right(rp4[1:].tolist(), int(e)), M4 - 1) for e in queries], np.int32)
    np.savez(HERE / "row_search.npz", rowptr=rp4, queries=queries, rows=rows)

    # 5. transpose: stable sort of the nnz by column
    order = np.argsort(ci, kind="stable")
    colptr = np.concatenate(([0], np.cumsum(np.bincount(ci, minlength=K)))).astype(np.int32)
    rowind = np.repeat(np.arange(M), deg)[order].astype(np.int32)
    np.savez(HERE / "transpose.npz", nrow=M, ncol=K, rowptr=rp, colind=ci, colptr=colptr, rowind=rowind,
             perm=order.astype(np.int32))


if __name__ == "__main__":
    main()
