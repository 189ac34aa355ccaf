"""Seeded synthetic sparse inputs (numpy, host side).

Sizes and laws follow the five workloads this library is measured on:
  * poisson_csr     16384 x 16384, nnz/row ~ Poisson(16), values U(-1, 1)
  * powerlaw_csr    262144 x 262144, Zipf(2.1) row degrees clipped to
                    [1, 20000] and scaled to ~4M nnz, values U(-1, 1)
  * banded_block_csr 120000 x 120000 of 3x3 dense blocks, ~20 nnz/row,
                    bandwidth 600, values N(0, 1)
  * gcn_normalized  D^-1/2 (A + I) D^-1/2 of a symmetric Chung-Lu power-law
                    graph, 131072 nodes, average degree 20
  * batched_graphs  256 graphs, node counts log-uniform in [64, 50000], Zipf(2.1)
                    degrees clipped at 2000, as one block-diagonal CSR
Columns are distinct and sorted within every row. All generators take a seed
and use numpy's PCG64 (np.random.default_rng).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class HostCsr:
    nrow: int
    ncol: int
    rowptr: np.ndarray  # int32 [nrow + 1]
    colind: np.ndarray  # int32 [nnz]
    vals: np.ndarray | None  # float32 [nnz]

    @property
    def nnz(self) -> int:
        return int(self.colind.size)

    def degrees(self) -> np.ndarray:
        return np.diff(self.rowptr.astype(np.int64))

    def row_ids(self) -> np.ndarray:
        return np.repeat(np.arange(self.nrow, dtype=np.int32), self.degrees())

    def to_device(self, device="cuda"):
        from .ops import CsrMatrix
        return CsrMatrix.from_numpy(self.nrow, self.ncol, self.rowptr, self.colind, self.vals, device=device)


def _distinct_sorted_columns(rng, deg: np.ndarray, K: int) -> np.ndarray:
    """Uniform columns without repetition inside each row, sorted per row."""
    deg = np.minimum(deg.astype(np.int64), K)
    nrow = deg.size
    row = np.repeat(np.arange(nrow, dtype=np.int64), deg)
    # dense rows (more than half of K) are drawn as permutations
    dense = np.nonzero(deg * 2 > K)[0]
    col = rng.integers(0, K, size=row.size, dtype=np.int64)
    starts = np.concatenate(([0], np.cumsum(deg)))
    for r in dense:
        col[starts[r]:starts[r + 1]] = np.sort(rng.permutation(K)[: deg[r]])
    for _ in range(64):
        key = row * K + col
        order = np.argsort(key, kind="stable")
        key = key[order]
        col = col[order]
        dup = np.zeros(key.size, dtype=bool)
        dup[1:] = key[1:] == key[:-1]
        if not dup.any():
            return col.astype(np.int32)
        col[dup] = rng.integers(0, K, size=int(dup.sum()), dtype=np.int64)
    raise RuntimeError("could not draw distinct columns")


def csr_from_degrees(rng, deg: np.ndarray, K: int, values: str = "uniform") -> HostCsr:
    deg = np.minimum(deg.astype(np.int64), K)
    rowptr = np.zeros(deg.size + 1, dtype=np.int64)
    np.cumsum(deg, out=rowptr[1:])
    if rowptr[-1] > np.iinfo(np.int32).max:
        raise ValueError("nnz exceeds int32")
    col = _distinct_sorted_columns(rng, deg, K)
    n = col.size
    if values == "uniform":
        vals = rng.uniform(-1.0, 1.0, size=n).astype(np.float32)
    elif values == "normal":
        vals = rng.standard_normal(n).astype(np.float32)
    elif values == "ones":
        vals = np.ones(n, dtype=np.float32)
    elif values == "int":
        vals = rng.integers(-2, 3, size=n).astype(np.float32)
    else:
        raise ValueError(values)
    return HostCsr(int(deg.size), int(K), rowptr.astype(np.int32), col, vals)


def poisson_csr(M: int = 16384, K: int = 16384, mean: float = 16.0, seed: int = 0, values="uniform") -> HostCsr:
    rng = np.random.default_rng(seed)
    deg = np.clip(rng.poisson(mean, size=M), 0, K)
    return csr_from_degrees(rng, deg, K, values)


def zipf_degrees(rng, n: int, alpha: float, max_deg: int, target_nnz: int | None) -> tuple[np.ndarray, float]:
    """Zipf(alpha) draws clipped to [1, max_deg]; with target_nnz the draws are
    scaled by s (bisection) so that sum(clip(round(s * z), 1, max_deg)) ~ target."""
    z = rng.zipf(alpha, size=n).astype(np.float64)
    if target_nnz is None:
        return np.clip(z, 1, max_deg).astype(np.int64), 1.0
    lo, hi = 0.1, 1000.0
    for _ in range(100):
        s = 0.5 * (lo + hi)
        tot = np.clip(np.rint(s * z), 1, max_deg).sum()
        if tot < target_nnz:
            lo = s
        else:
            hi = s
    s = 0.5 * (lo + hi)
    return np.clip(np.rint(s * z), 1, max_deg).astype(np.int64), s


def powerlaw_csr(M: int = 262144, K: int = 262144, alpha: float = 2.1, max_deg: int = 20000,
                 target_nnz: int | None = 4194304, seed: int = 0, values="uniform") -> HostCsr:
    rng = np.random.default_rng(seed)
    deg, _ = zipf_degrees(rng, M, alpha, max_deg, target_nnz)
    return csr_from_degrees(rng, deg, K, values)


def segment_index_from_degrees(deg: np.ndarray) -> np.ndarray:
    """Sorted COO index with run lengths deg (segment s repeated deg[s] times)."""
    return np.repeat(np.arange(deg.size, dtype=np.int32), deg.astype(np.int64))


def banded_block_csr(n_block_rows: int = 40000, blocks_per_row: int = 7, half_band_blocks: int = 100,
                     seed: int = 0) -> HostCsr:
    """3x3 dense blocks; every block row has its diagonal block plus
    blocks_per_row - 1 distinct off-diagonal block columns within
    |delta| <= half_band_blocks (bandwidth 2 * 3 * half_band_blocks = 600)."""
    rng = np.random.default_rng(seed)
    nb = n_block_rows
    rows_cols = []
    for br in range(nb):
        lo, hi = max(0, br - half_band_blocks), min(nb - 1, br + half_band_blocks)
        cand = np.arange(lo, hi + 1)
        cand = cand[cand != br]
        pick = rng.choice(cand, size=min(blocks_per_row - 1, cand.size), replace=False)
        rows_cols.append(np.sort(np.concatenate(([br], pick))))
    counts = np.array([c.size for c in rows_cols], dtype=np.int64)
    bcols = np.concatenate(rows_cols)
    # expand blocks: row 3*br + i has columns 3*bc + j for every block column bc
    M = 3 * nb
    deg = np.repeat(counts * 3, 3)
    rowptr = np.zeros(M + 1, dtype=np.int64)
    np.cumsum(deg, out=rowptr[1:])
    col = np.empty(rowptr[-1], dtype=np.int32)
    bstart = np.concatenate(([0], np.cumsum(counts)))
    for br in range(nb):
        bc = bcols[bstart[br]:bstart[br + 1]]
        cols = (3 * bc[:, None] + np.arange(3)[None, :]).reshape(-1).astype(np.int32)
        for i in range(3):
            r = 3 * br + i
            col[rowptr[r]:rowptr[r + 1]] = cols
    vals = rng.standard_normal(col.size).astype(np.float32)
    return HostCsr(M, M, rowptr.astype(np.int32), col, vals)


def gcn_normalized(n: int = 131072, avg_deg: float = 20.0, alpha: float = 2.1, seed: int = 0) -> HostCsr:
    """Symmetric Chung-Lu power-law graph plus self loops, normalized
    D^-1/2 (A + I) D^-1/2 in float64 and stored as float32."""
    rng = np.random.default_rng(seed)
    w = rng.zipf(alpha, size=n).astype(np.float64)
    w = np.minimum(w, np.sqrt(n * avg_deg))
    m_half = int(n * avg_deg / 2)
    p = w / w.sum()
    src = rng.choice(n, size=m_half, p=p)
    dst = rng.choice(n, size=m_half, p=p)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    a = np.concatenate([src, dst, np.arange(n)])
    b = np.concatenate([dst, src, np.arange(n)])
    key = np.unique(a.astype(np.int64) * n + b)
    r = (key // n).astype(np.int64)
    c = (key % n).astype(np.int32)
    deg = np.bincount(r, minlength=n).astype(np.float64)
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(r, minlength=n), out=rowptr[1:])
    dinv = 1.0 / np.sqrt(deg)
    vals = (dinv[r] * dinv[c]).astype(np.float32)
    return HostCsr(n, n, rowptr.astype(np.int32), c, vals)


def batched_graphs(n_graphs: int = 256, min_nodes: int = 64, max_nodes: int = 50000, alpha: float = 2.1,
                   max_deg: int = 2000, seed: int = 0) -> tuple[HostCsr, np.ndarray]:
    """Block-diagonal CSR of n_graphs graphs; returns it and the node offsets
    (graph g owns rows/columns [offsets[g], offsets[g+1]))."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(min_nodes), np.log(max_nodes), size=n_graphs)).astype(np.int64)
    offsets = np.zeros(n_graphs + 1, dtype=np.int64)
    np.cumsum(sizes, out=offsets[1:])
    rowptrs, cols = [np.zeros(1, dtype=np.int64)], []
    total = 0
    for g in range(n_graphs):
        ng = int(sizes[g])
        deg = np.clip(rng.zipf(alpha, size=ng), 1, min(max_deg, ng)).astype(np.int64)
        sub = csr_from_degrees(rng, deg, ng, values="ones")
        rowptrs.append(sub.rowptr[1:].astype(np.int64) + total)
        cols.append(sub.colind.astype(np.int64) + offsets[g])
        total += sub.nnz
    rowptr = np.concatenate(rowptrs)
    col = np.concatenate(cols).astype(np.int32)
    vals = rng.uniform(-1.0, 1.0, size=col.size).astype(np.float32)
    M = int(offsets[-1])
    return HostCsr(M, M, rowptr.astype(np.int32), col, vals), offsets


def small_random_csr(M: int, K: int, density: float, seed: int, values="uniform", empty_rows: float = 0.0) -> HostCsr:
    """Small matrices for parity tests (optionally with forced empty rows)."""
    rng = np.random.default_rng(seed)
    deg = rng.binomial(K, density, size=M)
    if empty_rows > 0:
        deg[rng.random(M) < empty_rows] = 0
    return csr_from_degrees(rng, deg, K, values)
