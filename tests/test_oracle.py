"""Pins the CPU oracle (oracle/oracle.cpp) before anything is checked against it.

Each restated reference function is checked against something that does not
share its code: the committed fixtures (dense float64 products, np.add.at,
bisect, stable argsort), brute force on random small inputs, closed forms (the
reference harness's all-ones known answer), and invariants. Planted mistakes
(a dropped term, a shifted column index, a transposed operand) must fail the
same checks, so the checks are not vacuous.
"""
import bisect
from pathlib import Path

import numpy as np
import pytest

from segscan_amd import synth

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name):
    return np.load(GOLDEN / name, allow_pickle=False)


def dense(m):
    D = np.zeros((m.nrow, m.ncol), np.float64)
    D[m.row_ids(), m.colind] = m.vals
    return D


# ---- golden fixtures -------------------------------------------------------------------------

def test_spmm_tiny_fixture(oracle):
    f = load("spmm_tiny.npz")
    C32 = oracle.spmm_f32(f["rowptr"], f["colind"], f["vals"], f["B"])
    np.testing.assert_allclose(C32, f["C"], rtol=1e-6, atol=1e-6)
    C64, _ = oracle.spmm_f64(f["rowptr"], f["colind"], f["vals"], f["B"])
    np.testing.assert_allclose(C64, f["C"], rtol=1e-15, atol=1e-15)
    assert np.all(C32[[0, 3, 4, 7]] == 0)  # empty rows stay exactly zero


def test_spmm_int_fixture(oracle):
    f = load("spmm_int.npz")
    C = oracle.spmm_f32(f["rowptr"], f["colind"], f["vals"], f["B"])
    assert np.array_equal(C, f["C"].astype(np.float32))


def test_segment_int_fixture(oracle):
    f = load("segment_int.npz")
    dst = oracle.segment_coo_f32(f["src"], f["index"], f["dst"].shape[0])
    assert np.array_equal(dst, f["dst"].astype(np.float32))


def test_row_search_fixture(oracle):
    f = load("row_search.npz")
    got = np.array([oracle.row_of_nnz(f["rowptr"], int(e)) for e in f["queries"]])
    assert np.array_equal(got, f["rows"])


def test_transpose_fixture(oracle):
    f = load("transpose.npz")
    colptr, rowind, perm = oracle.transpose(int(f["nrow"]), int(f["ncol"]), f["rowptr"], f["colind"])
    assert np.array_equal(colptr, f["colptr"])
    assert np.array_equal(rowind, f["rowind"])
    assert np.array_equal(perm, f["perm"])


# ---- brute force, closed forms, invariants ---------------------------------------------------

@pytest.mark.parametrize("seed", range(5))
def test_row_search_brute_force(oracle, seed):
    rng = np.random.default_rng(seed)
    deg = rng.integers(0, 4, size=rng.integers(1, 40))
    deg[rng.random(deg.size) < 0.3] = 0
    rp = np.concatenate(([0], np.cumsum(deg))).astype(np.int32)
    M = deg.size
    for e in range(int(rp[-1]) + 2):
        expect = min(bisect.bisect_right(rp[1:].tolist(), e), M - 1)
        assert oracle.row_of_nnz(rp, e) == expect
        if e < rp[-1]:
            assert rp[expect] <= e < rp[expect + 1]


@pytest.mark.parametrize("seed", range(4))
def test_spmm_dense_brute_force(oracle, seed):
    m = synth.small_random_csr(40, 30, 0.2, seed=seed, empty_rows=0.2)
    B = np.random.default_rng(seed).uniform(-1, 1, size=(30, 9)).astype(np.float32)
    ref = dense(m) @ B.astype(np.float64)
    C64, D = oracle.spmm_f64(m.rowptr, m.colind, m.vals, B)
    np.testing.assert_allclose(C64, ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(D, np.abs(dense(m)) @ np.abs(B.astype(np.float64)), rtol=1e-12, atol=1e-12)
    C32 = oracle.spmm_f32(m.rowptr, m.colind, m.vals, B)
    assert oracle.normalized_error(C32, C64, D) <= 1e-6


def test_spmm_sequential_order_is_the_reference_loop(oracle):
    """fp32 rounding follows the i -> p -> j order of spmm_reference_host: a
    float32 numpy loop in that order gives the same bits."""
    m = synth.small_random_csr(30, 25, 0.3, seed=9)
    B = np.random.default_rng(9).uniform(-1, 1, size=(25, 6)).astype(np.float32)
    C = np.zeros((30, 6), np.float32)
    for i in range(30):
        for p in range(m.rowptr[i], m.rowptr[i + 1]):
            C[i] = (C[i] + np.float32(m.vals[p]) * B[m.colind[p]]).astype(np.float32)
    assert np.array_equal(oracle.spmm_f32(m.rowptr, m.colind, m.vals, B), C)


def test_spmm_all_ones_known_answer(oracle):
    """test.cu:42 and dataloader.hpp:154-155: values and B all one -> C[i][j] = deg(i)."""
    m = synth.powerlaw_csr(M=3000, K=3000, max_deg=500, target_nnz=30000, seed=1)
    B = np.ones((3000, 16), np.float32)
    C = oracle.spmm_f32(m.rowptr, m.colind, None, B)
    assert np.array_equal(C, np.repeat(m.degrees()[:, None].astype(np.float32), 16, axis=1))


def test_spmm_linearity(oracle):
    m = synth.small_random_csr(50, 40, 0.15, seed=3, values="int")
    rng = np.random.default_rng(3)
    B1 = rng.integers(-3, 4, size=(40, 5)).astype(np.float32)
    B2 = rng.integers(-3, 4, size=(40, 5)).astype(np.float32)
    lhs = oracle.spmm_f32(m.rowptr, m.colind, m.vals, 2 * B1 - B2)
    rhs = 2 * oracle.spmm_f32(m.rowptr, m.colind, m.vals, B1) - oracle.spmm_f32(m.rowptr, m.colind, m.vals, B2)
    assert np.array_equal(lhs, rhs)


def test_spmm_accumulates_into_c(oracle):
    m = synth.small_random_csr(20, 20, 0.2, seed=4, values="int")
    B = np.random.default_rng(4).integers(-2, 3, size=(20, 3)).astype(np.float32)
    C0 = np.full((20, 3), 5.0, np.float32)
    C = oracle.spmm_f32(m.rowptr, m.colind, m.vals, B, C=C0.copy())
    assert np.array_equal(C, C0 + oracle.spmm_f32(m.rowptr, m.colind, m.vals, B))


@pytest.mark.parametrize("seed", range(3))
def test_segment_brute_force(oracle, seed):
    rng = np.random.default_rng(seed)
    deg = rng.integers(0, 6, size=30)
    idx = synth.segment_index_from_degrees(deg)
    src = rng.uniform(-1, 1, size=(idx.size, 4)).astype(np.float32)
    ref = np.zeros((30, 4), np.float64)
    np.add.at(ref, idx, src.astype(np.float64))
    dst64, absd = oracle.segment_coo_f64(src, idx, 30)
    np.testing.assert_allclose(dst64, ref, rtol=1e-12, atol=1e-12)
    assert oracle.normalized_error(oracle.segment_coo_f32(src, idx, 30), dst64, absd) <= 1e-6
    # closed forms: src = 1 gives the run lengths, src[e] = e gives index sums
    ones = oracle.segment_coo_f32(np.ones((idx.size, 1), np.float32), idx, 30)
    assert np.array_equal(ones[:, 0], deg.astype(np.float32))
    ar = oracle.segment_coo_f32(np.arange(idx.size, dtype=np.float32)[:, None], idx, 30)
    assert np.array_equal(ar[:, 0], np.bincount(idx, weights=np.arange(idx.size), minlength=30).astype(np.float32))


@pytest.mark.parametrize("seed", range(3))
def test_transpose_stable_sort(oracle, seed):
    m = synth.small_random_csr(60, 45, 0.1, seed=seed, empty_rows=0.2)
    colptr, rowind, perm = oracle.transpose(m.nrow, m.ncol, m.rowptr, m.colind)
    order = np.argsort(m.colind, kind="stable")
    assert np.array_equal(perm, order)
    assert np.array_equal(rowind, m.row_ids()[order])
    assert np.array_equal(colptr, np.concatenate(([0], np.cumsum(np.bincount(m.colind, minlength=45)))))
    # A^T A^T = A: transposing twice restores the CSR arrays
    c2, r2, p2 = oracle.transpose(m.ncol, m.nrow, colptr, rowind)
    assert np.array_equal(c2, m.rowptr) and np.array_equal(r2, m.colind)


def test_sddmm_brute_force(oracle):
    m = synth.small_random_csr(30, 20, 0.25, seed=6)
    rng = np.random.default_rng(6)
    X = rng.standard_normal((30, 7)).astype(np.float32)
    Y = rng.standard_normal((20, 7)).astype(np.float32)
    out, absd = oracle.sddmm_f64(m.rowptr, m.colind, X, Y)
    ref = np.einsum("pj,pj->p", X[m.row_ids()].astype(np.float64), Y[m.colind].astype(np.float64))
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(oracle.sddmm_f64d(m.rowptr, m.colind, X, Y), ref, rtol=1e-12, atol=1e-12)


def test_check_result_rule(oracle):
    """check.cuh:116-136: |c - ref| > 1e-2 |ref| fails, counted per row."""
    ref = np.array([[1.0, 2.0], [0.0, 4.0]], np.float32)
    assert oracle.check_result(ref, ref) == 0
    assert oracle.check_result(ref * 1.009, ref) == 0
    assert oracle.check_result(ref * 1.011, ref) == 2  # both rows fail (row 1 on 4.0; 0 * 1.011 is exact)
    bad = ref.copy()
    bad[1, 0] = 1e-6  # non-zero against a zero reference always fails
    assert oracle.check_result(bad, ref) == 1


@pytest.mark.parametrize("cv", [0.0, 0.5, 2.0])
def test_generate_index_properties(oracle, cv):
    """dataloader.h:33-67: sorted runs, per-run clamp, the last id of the range
    absorbs the remainder, early stop once the total is reached."""
    idx, dst_len = oracle.generate_index(100, 2, 80, 5000, cv)
    assert idx.size == 5000
    assert np.all(np.diff(idx) >= 0) and idx.min() >= 0 and idx.max() < 100
    runs = np.bincount(idx, minlength=100)
    assert dst_len == np.count_nonzero(runs)
    last = idx.max()
    body = runs[:last]
    assert np.all((body >= 2) & (body <= 80))


# ---- planted mistakes must be caught ---------------------------------------------------------

def test_planted_mistakes_are_detected(oracle):
    f = load("spmm_int.npz")
    rp, ci, va, B, C = f["rowptr"], f["colind"], f["vals"], f["B"], f["C"].astype(np.float32)
    # dropped last term of every row
    keep = np.ones(ci.size, bool)
    keep[rp[1:][np.diff(rp) > 0] - 1] = False
    counts = np.array([keep[rp[i]:rp[i + 1]].sum() for i in range(rp.size - 1)])
    rp_drop = np.concatenate(([0], np.cumsum(counts))).astype(np.int32)
    assert not np.array_equal(oracle.spmm_f32(rp_drop, ci[keep], va[keep], B), C)
    # shifted column index
    assert not np.array_equal(oracle.spmm_f32(rp, (ci + 1) % B.shape[0], va, B), C)
    # wrong sign
    assert not np.array_equal(oracle.spmm_f32(rp, ci, -va, B), C)
    # transposed operand: a square case where A^T B differs from A B
    m = synth.small_random_csr(20, 20, 0.2, seed=12, values="int")
    Bs = np.random.default_rng(12).integers(-2, 3, size=(20, 4)).astype(np.float32)
    At = oracle.transpose(20, 20, m.rowptr, m.colind)
    assert not np.array_equal(oracle.spmm_f32(At[0], At[1], m.vals[At[2]], Bs), dense(m) @ Bs)
