"""Tuning tables in the C ABI (segscan_tune_table_*, segscan_autotune_spmm): the keys
the C++ lookup builds equal the Python autotuner's (same JSON files), bucket knobs
resolve the same way, entries of another architecture or ROCm release are dropped on
load, save / load round-trips; on the GPU the C++ autotuner's winner is correct against
the oracle and is recorded under the matrix's key."""
import json

import numpy as np
import pytest
import torch

from segscan_amd import _abi, autotune, ops, synth


def _host_matrix(rows, ncol, nnz, max_row):
    """A descriptor-only matrix (CPU tensors; lookup reads sizes only)."""
    rp = torch.zeros(rows + 1, dtype=torch.int32)
    ci = torch.zeros(max(nnz, 1), dtype=torch.int32)[:nnz]
    return ops.CsrMatrix(rows, ncol, rp, ci, None, max_row)


def _stats(rows, nnz, max_row):
    return {"rows": rows, "nnz": nnz, "max": max_row, "mean": nnz / max(rows, 1)}


def test_prefix_names_arch_and_rocm():
    pre = autotune.tune_prefix()
    arch, rocm = pre.split("|")
    assert arch == "gfx950" and rocm.startswith("rocm") and "." in rocm


@pytest.mark.parametrize("rows,nnz,max_row", [
    (262144, 4194295, 20000), (1000, 16000, 80), (1000, 16000, 96), (1000, 16000, 97),  # even / skew boundary
    (4096, 65536, 256), (4096, 65536, 4096), (100, 100, 1600), (5, 1 << 20, 600000), (3000, 2999, 7)])
@pytest.mark.parametrize("N,dtype", [(64, torch.float32), (100, torch.bfloat16), (3, torch.float16)])
def test_c_lookup_finds_python_keys(tmp_path, rows, nnz, max_row, N, dtype):
    st = _stats(rows, nnz, max_row)
    bucket = autotune.table_key("spmm", dtype, N, st)
    mkey = autotune.matrix_key("spmm", dtype, N, st, 777)
    knobs_b = {"coarsen_factor": 2, "group_factor": 5, "thread_nz": 16, "thread_per_block": 256, "waves_per_cu": 8,
               "write_policy": 2}
    path = tmp_path / "t.json"
    path.write_text(json.dumps({bucket: {"variant": "ebsr", "knobs": knobs_b, "action": "x", "matrices": ["a", "b"]}}))
    tab = autotune.NativeTable(path)
    assert (tab.loaded, tab.dropped) == (1, 0)
    v, k, found = tab.lookup(_host_matrix(rows, 777, nnz, max_row), N, dtype)
    assert found == 1 and v == _abi.SPMM_EBSR
    want = autotune.resolve_knobs(knobs_b, st)
    assert {n: x for n, x in k.items() if x} == want
    # the matrix's own entry wins over the bucket
    path.write_text(json.dumps({bucket: {"variant": "ebsr", "knobs": knobs_b},
                                mkey: {"variant": "rbpr", "knobs": {"group_factor": 3, "tile_factor": 3}, "ms": 1.0}}))
    tab = autotune.NativeTable(path)
    v, k, found = tab.lookup(_host_matrix(rows, 777, nnz, max_row), N, dtype)
    assert found == 2 and v == _abi.SPMM_RBPR and k["group_factor"] == 3 and k["tile_factor"] == 3
    # another N bucket or dtype finds nothing: AUTO
    v, k, found = tab.lookup(_host_matrix(rows, 777, nnz, max_row), 2 * N + 200, dtype)
    assert found == 0 and v == _abi.SPMM_AUTO


def test_stale_entries_dropped_and_round_trip(tmp_path):
    cur = autotune.tune_prefix()
    entries = {
        f"{cur}|spmm|f32|N64|skew0|s": {"variant": "rbsr", "knobs": {"coarsen_factor": 4}},
        f"{cur}|spmm|bf16|N64|skew3|l|m10x10:50:20": {"knobs": None, "ms": 0.5, "variant": "auto"},
        "gfx950|rocm6.1|spmm|f32|N64|skew0|s": {"variant": "ebpr", "knobs": None},  # older ROCm
        "gfx950|spmm|f32|N64|skew0|s": {"variant": "ebpr", "knobs": None},          # key without a release
        "gfx942|rocm7.2|spmm|f32|N64|skew0|s": {"variant": "ebpr", "knobs": None},  # another GPU
    }
    src = tmp_path / "in.json"
    src.write_text(json.dumps(entries, indent=1))
    tab = autotune.NativeTable(src)
    assert (tab.loaded, tab.dropped) == (2, 3)
    stale = []
    py = autotune.load_table(src, stale=stale)
    assert len(stale) == 3 and set(py) == {k for k in entries if k.startswith(cur + "|")}
    out = tmp_path / "out.json"
    tab.save(out)
    assert json.loads(out.read_text()) == py


def test_malformed_tables_rejected(tmp_path, ):
    lib = _abi.load()
    tab = autotune.NativeTable()
    for text in ('[1, 2]', '{"k": {"variant": "nope"}}', '{"' + autotune.tune_prefix() + '|x": {"variant": 3}}',
                 '{"a": 1', ''):
        p = tmp_path / "bad.json"
        p.write_text(text)
        if text.startswith('{"k"'):
            tab.load(p)  # a key of another prefix is dropped before its value is read
            assert tab.dropped == 1
            continue
        with pytest.raises(_abi.SegscanError) as exc:
            tab.load(p)
        assert exc.value.status == _abi.EINVAL
    assert lib.segscan_tune_table_load(None, b"x", None, None) == _abi.EINVAL
    with pytest.raises(_abi.SegscanError):
        tab.load(tmp_path / "missing.json")


@pytest.mark.gpu
def test_native_autotune_records_a_correct_winner(gpu, oracle, tmp_path):
    m = synth.powerlaw_csr(M=20000, K=20000, max_deg=3000, target_nnz=300_000, seed=11)
    A = m.to_device(gpu)
    rng = np.random.default_rng(12)
    Bh = rng.uniform(-1, 1, size=(m.ncol, 64)).astype(np.float32)
    B = torch.from_numpy(Bh).to(gpu)
    tab = autotune.NativeTable()
    v, k, ms = tab.autotune(A, B, reps=3)
    print(f"native autotune: variant {v} knobs {k} {ms:.4f} ms", flush=True)
    assert 0 < ms < 50
    C = ops.spmm(A, B, variant=v, knobs={n: x for n, x in k.items() if x} or None)
    C64, D = oracle.spmm_f64(m.rowptr, m.colind, m.vals, Bh)
    assert oracle.normalized_error(C.cpu().numpy(), C64, D) <= 1e-6
    v2, k2, found = tab.lookup(A, 64, torch.float32)
    assert found == 2 and v2 == v and k2 == k
    out = tmp_path / "tuned.json"
    tab.save(out)
    py = autotune.load_table(out)
    key = autotune.matrix_key("spmm", torch.float32, 64, autotune.row_stats(m.rowptr), m.ncol)
    assert key in py and py[key]["ms"] == pytest.approx(ms, rel=1e-3)
