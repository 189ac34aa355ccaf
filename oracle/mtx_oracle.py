"""Pure-Python restatement of the reference's Matrix Market reader — TEST
INFRASTRUCTURE ONLY (the checker of segscan_mtx_read; the product never imports it).

read_mtx_file, test-cg/include/dataloader/dataloader.hpp:22-107, with the NIST
helpers it calls (mmio.hpp:203 mm_read_banner, :285 mm_read_mtx_crd_size):
  * entries "row col [value [imag]]", 1-based → 0-based (:48-64);
  * "symmetric" (mm_is_symmetric): each off-diagonal (i, j) also yields (j, i),
    then sort and drop exact duplicates (:69-86); otherwise sort only (:87-89);
  * CSR built by walking the sorted list while curr_pos < the DECLARED count
    (:92-106), so a mirrored list is cut to that count (``quirks=True``).
``quirks=False`` is the corrected reader: skew-symmetric (negated) and hermitian
(real part) files mirrored too, no cut.

Pinned by tests/test_mtx.py: the corrected reader against scipy.io.mmread,
the quirks against hand-worked files.
"""
from __future__ import annotations

import numpy as np


def read_mtx(path, keep_values: bool = False, quirks: bool = False):
    with open(path) as f:
        lines = f.read().split("\n")
    head = lines[0].split()
    assert head[0] == "%%MatrixMarket" and head[1].lower() == "matrix" and head[2].lower() == "coordinate"
    field, sym = head[3].lower(), head[4].lower()
    body = [ln for ln in lines[1:] if ln.strip() and not ln.startswith("%")]
    nrow, ncol, nnz = (int(t) for t in body[0].split())
    toks = " ".join(body[1:]).split()
    width = {"pattern": 2, "complex": 4}.get(field, 3)
    mirror = sym == "symmetric" or (not quirks and sym in ("skew-symmetric", "hermitian"))
    coords = []
    for e in range(nnz):
        t = toks[e * width:(e + 1) * width]
        r, c = int(t[0]) - 1, int(t[1]) - 1
        v = float(t[2]) if (keep_values and width > 2) else 1.0
        coords.append((r, c, v))
        if mirror and r != c:
            coords.append((c, r, -v if (keep_values and sym == "skew-symmetric") else v))
    coords.sort(key=lambda x: (x[0], x[1]))  # stable, like the C++ reader's stable_sort
    if mirror:
        kept = []
        for x in coords:
            if not kept or (kept[-1][0], kept[-1][1]) != (x[0], x[1]):
                kept.append(x)
        coords = kept
    if quirks:
        coords = coords[:nnz]
    rowptr = np.zeros(nrow + 1, dtype=np.int32)
    for r, _, _ in coords:
        rowptr[r + 1] += 1
    rowptr = np.cumsum(rowptr).astype(np.int32)
    colind = np.array([c for _, c, _ in coords], dtype=np.int32)
    vals = np.array([v for _, _, v in coords], dtype=np.float32)
    return nrow, ncol, rowptr, colind, vals
