"""Matrix Market reader (segscan_mtx_read) — the reference's read_mtx_file,
test-cg/include/dataloader/dataloader.hpp:22-107.

The pure-Python restatement (oracle/mtx_oracle.py) is pinned two ways: its
corrected mode against scipy.io.mmread, which implements the Matrix Market
format independently, and its reference-quirk mode against hand-worked files.
The library's reader is then checked against the restatement, bit-exact.
"""
import zlib

import numpy as np
import pytest
import scipy.io

from oracle import mtx_oracle
from segscan_amd import mtx


def _write(path, field, sym, nrow, ncol, entries, comments=("% written by the test",)):
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {field} {sym}\n")
        for c in comments:
            f.write(c + "\n")
        f.write(f"{nrow} {ncol} {len(entries)}\n")
        for e in entries:
            f.write(" ".join(str(x) for x in e) + "\n")
    return path


def _random_entries(rng, nrow, ncol, n, field, sym, dups=False):
    ents = set()
    out = []
    while len(out) < n:
        r, c = int(rng.integers(1, nrow + 1)), int(rng.integers(1, ncol + 1))
        if sym != "general":
            r, c = max(r, c), min(r, c)
            if sym == "skew-symmetric" and r == c:
                continue
        if (r, c) in ents and not dups:
            continue
        ents.add((r, c))
        if field == "pattern":
            out.append((r, c))
        elif field == "integer":
            out.append((r, c, int(rng.integers(-9, 10))))
        elif field == "complex":
            out.append((r, c, round(float(rng.normal()), 6), round(float(rng.normal()), 6)))
        else:
            out.append((r, c, round(float(rng.normal()), 6)))
    return out


def _triples(rowptr, colind, vals):
    rows = np.repeat(np.arange(rowptr.size - 1), np.diff(rowptr))
    order = np.lexsort((vals, colind, rows))
    return rows[order], colind[order], vals[order]


def _as_csr_triples(nrow, ncol, rowptr, colind, vals):
    assert rowptr.shape == (nrow + 1,) and rowptr[0] == 0 and rowptr[-1] == colind.size
    assert np.all(np.diff(rowptr) >= 0)
    return _triples(rowptr, colind, vals)


CASES = [
    ("real", "general"), ("integer", "general"), ("pattern", "general"),
    ("real", "symmetric"), ("pattern", "symmetric"), ("real", "skew-symmetric"),
    ("complex", "hermitian"), ("complex", "general"),
]


@pytest.mark.parametrize("field,sym", CASES)
def test_oracle_matches_scipy(tmp_path, field, sym):
    rng = np.random.default_rng(zlib.crc32(f"{field}/{sym}".encode()))
    nrow = ncol = 40 if sym != "general" else None
    if nrow is None:
        nrow, ncol = 37, 53
    ents = _random_entries(rng, nrow, ncol, 200, field, sym, dups=(sym == "general"))
    p = _write(tmp_path / "m.mtx", field, sym, nrow, ncol, ents)
    n, k, rp, ci, va = mtx_oracle.read_mtx(p, keep_values=True)
    ref = scipy.io.mmread(str(p)).tocoo()
    assert (n, k) == ref.shape
    data = np.real(ref.data) if np.iscomplexobj(ref.data) else ref.data
    order = np.lexsort((data, ref.col, ref.row))
    got = _as_csr_triples(n, k, rp, ci, va)
    np.testing.assert_array_equal(got[0], ref.row[order])
    np.testing.assert_array_equal(got[1], ref.col[order])
    np.testing.assert_allclose(got[2], data[order].astype(np.float32), rtol=0, atol=0)


def test_oracle_reference_quirks_hand_worked(tmp_path):
    # lower triangle of [[2,3,0],[3,0,4],[0,4,0]]
    p = _write(tmp_path / "s.mtx", "real", "symmetric", 3, 3, [(1, 1, 2.0), (2, 1, 3.0), (3, 2, 4.0)])
    _, _, rp, ci, va = mtx_oracle.read_mtx(p, keep_values=True)
    assert rp.tolist() == [0, 2, 4, 5] and ci.tolist() == [0, 1, 0, 2, 1] and va.tolist() == [2, 3, 3, 4, 4]
    # read_mtx_file stops at the declared count (3): rows 1 and 2 lose entries
    _, _, rp, ci, va = mtx_oracle.read_mtx(p, quirks=True)
    assert rp.tolist() == [0, 2, 3, 3] and ci.tolist() == [0, 1, 0] and va.tolist() == [1, 1, 1]
    # skew-symmetric is not mirrored by read_mtx_file (mm_is_symmetric is false)
    p = _write(tmp_path / "k.mtx", "real", "skew-symmetric", 3, 3, [(2, 1, 5.0), (3, 1, -1.0)])
    _, _, rp, ci, _ = mtx_oracle.read_mtx(p, quirks=True)
    assert rp.tolist() == [0, 0, 1, 2] and ci.tolist() == [0, 0]
    _, _, rp, ci, va = mtx_oracle.read_mtx(p, keep_values=True)
    assert rp.tolist() == [0, 2, 3, 4] and ci.tolist() == [1, 2, 0, 0] and va.tolist() == [-5, 1, 5, -1]
    # general files: sorted, duplicates kept, no cut
    p = _write(tmp_path / "g.mtx", "pattern", "general", 2, 3, [(2, 3), (1, 2), (2, 3), (1, 1)])
    _, _, rp, ci, _ = mtx_oracle.read_mtx(p, quirks=True)
    assert rp.tolist() == [0, 2, 4] and ci.tolist() == [0, 1, 2, 2]


@pytest.mark.parametrize("field,sym", CASES)
@pytest.mark.parametrize("keep_values", [False, True])
@pytest.mark.parametrize("quirks", [False, True])
def test_library_matches_oracle(tmp_path, field, sym, keep_values, quirks):
    rng = np.random.default_rng(7 + len(field) * 13 + len(sym))
    nrow, ncol = (64, 64) if sym != "general" else (45, 70)
    ents = _random_entries(rng, nrow, ncol, 300, field, sym, dups=(sym == "general"))
    p = _write(tmp_path / "m.mtx", field, sym, nrow, ncol, ents, comments=("%", "% two comment lines"))
    n, k, rp, ci, va = mtx_oracle.read_mtx(p, keep_values=keep_values, quirks=quirks)
    got = mtx.read_mtx(p, keep_values=keep_values, reference_quirks=quirks)
    assert (got.nrow, got.ncol) == (n, k)
    np.testing.assert_array_equal(got.rowptr, rp)
    np.testing.assert_array_equal(got.colind, ci)
    np.testing.assert_array_equal(got.vals, va)


def test_library_empty_and_edge_shapes(tmp_path):
    p = _write(tmp_path / "e.mtx", "real", "general", 5, 4, [])
    m = mtx.read_mtx(p)
    assert m.nnz == 0 and m.rowptr.tolist() == [0] * 6
    p = _write(tmp_path / "z.mtx", "pattern", "symmetric", 0, 0, [])
    m = mtx.read_mtx(p)
    assert m.nrow == 0 and m.rowptr.tolist() == [0]
    # Windows line ends and blank lines before the size line
    p = tmp_path / "w.mtx"
    p.write_bytes(b"%%MatrixMarket matrix coordinate real general\r\n\r\n% c\r\n2 2 2\r\n1 2 0.5\r\n2 1 -1.5\r\n")
    m = mtx.read_mtx(p, keep_values=True)
    assert m.rowptr.tolist() == [0, 1, 2] and m.colind.tolist() == [1, 0] and m.vals.tolist() == [0.5, -1.5]


@pytest.mark.parametrize("text,msg", [
    ("", "empty"),
    ("%%NotMatrixMarket matrix coordinate real general\n1 1 0\n", "banner"),
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", "coordinate"),
    ("%%MatrixMarket matrix coordinate quaternion general\n1 1 0\n", "field"),
    ("%%MatrixMarket matrix coordinate real upper\n1 1 0\n", "symmetry"),
    ("%%MatrixMarket matrix coordinate real general\n2 2\n", "size"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n", "not enough"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n", "out of range"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n0 1 1.0\n", "out of range"),
])
def test_library_rejects_malformed(tmp_path, text, msg):
    p = tmp_path / "bad.mtx"
    p.write_text(text)
    with pytest.raises(mtx.MtxError, match=msg):
        mtx.read_mtx(p)


def test_library_missing_file(tmp_path):
    with pytest.raises(mtx.MtxError, match="cannot open"):
        mtx.read_mtx(tmp_path / "nope.mtx")


@pytest.mark.gpu
def test_mtx_to_spmm(tmp_path, gpu, oracle):
    """A file read by the library goes straight into the SpMM (values forced to one,
    as the reference's SpMatCsrDescr_t does, dataloader.hpp:154-155)."""
    import torch
    from segscan_amd import ops
    rng = np.random.default_rng(5)
    ents = _random_entries(rng, 300, 300, 2000, "real", "symmetric")
    p = _write(tmp_path / "s.mtx", "real", "symmetric", 300, 300, ents)
    h = mtx.read_mtx(p)
    B = rng.uniform(-1, 1, size=(300, 32)).astype(np.float32)
    C = ops.spmm(h.to_device(gpu), torch.from_numpy(B).to(gpu), variant="auto")
    C64, D = oracle.spmm_f64(h.rowptr, h.colind, h.vals, B)
    assert oracle.normalized_error(C.cpu().numpy(), C64, D) <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["", "quirks"])
def test_cpp_check_mode_on_mtx(tmp_path, gpu, mode):
    """`spmm_check <m.mtx> <N>`: the reference's check mode (test.cu:23-82) through
    SingleDataLoader and every launcher, plus DataLoader's host transpose checked
    against the device transpose."""
    import subprocess
    from pathlib import Path
    exe = Path(__file__).resolve().parent / "cpp" / "spmm_check"
    if not exe.exists():
        pytest.fail("tests/cpp/spmm_check is not built (run __graft_entry__.build())")
    rng = np.random.default_rng(21)
    p = _write(tmp_path / "s.mtx", "real", "symmetric", 700, 700,
               _random_entries(rng, 700, 700, 6000, "real", "symmetric"))
    res = subprocess.run([str(exe), str(p), "48", *([mode] if mode else [])], capture_output=True, text=True,
                         timeout=120)
    print(res.stdout[-3000:])
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    assert "Not Passed!" not in res.stdout and res.stdout.count("Passed!") >= 14
