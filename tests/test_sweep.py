"""Sweep / clean / report pipeline (segscan_amd.sweep): the analyses of the
reference's count.py, static.py and dgcmp.py (test-cg/src) on hand-worked
tables, the rocprofv3 CSV join on synthetic profiler files, and a tiny GPU run."""
import json
import math

import numpy as np
import pytest
import scipy.stats

from segscan_amd import sweep


def _rows(table, N=64, dtype="f32"):
    fam = {"plain_ebpr": ("ebpr", "plain"), "ebpr_default": ("ebpr", "default"),
           "a1": ("ebpr", "cg"), "a2": ("ebpr", "cg"), "hipsparse_default": ("vendor", "vendor"),
           "auto": ("select", "auto")}
    out = []
    for ds, times in table.items():
        for act, t in times.items():
            f, k = fam[act]
            out.append({"dataset": ds, "N": N, "dtype": dtype, "action": act, "family": f, "kind": k, "ms": t})
    return out


TABLE = {
    "d1": {"plain_ebpr": 4.0, "ebpr_default": 2.0, "a1": 1.0, "a2": 1.2, "hipsparse_default": 3.0, "auto": 1.5},
    "d2": {"plain_ebpr": 9.0, "ebpr_default": 3.3, "a1": 6.0, "a2": 3.0, "hipsparse_default": 1.5, "auto": 3.0},
}


def test_gmean_matches_scipy():
    x = np.random.default_rng(0).uniform(0.5, 3, size=50)
    assert sweep.gmean(x) == pytest.approx(scipy.stats.gmean(x), rel=1e-12)


def test_analyse_hand_worked():
    res = sweep.analyse(_rows(TABLE))["N64_f32"]
    fe = res["families"]["ebpr"]
    g = fe["gain_vs_plain"]  # ratios 4 and 3
    assert (g["mean"], g["std"], g["max"], g["argmax"], g["above_1"]) == (3.5, 0.5, 4.0, "d1", 2)
    assert g["gmean"] == pytest.approx(math.sqrt(12))
    g = fe["gain_vs_default"]  # 2 and 1.1
    assert g["mean"] == pytest.approx(1.55) and g["gmean"] == pytest.approx(math.sqrt(2.2))
    v = fe["vs_vendor"]  # 3 and 0.5: faster than the vendor on one matrix
    assert v["above_1"] == 1 and v["above_1_names"] == ["d1"] and v["gmean"] == pytest.approx(math.sqrt(1.5))
    # static gmeans: plain 3.46, default 1.48, a1 1.41, a2 1.095
    assert fe["best_static"] == "a2" and fe["static_ties"] == ["a2"]
    d = fe["dynamic_vs_static"]  # 1.2 and 1.0
    assert d["mean"] == pytest.approx(1.1) and d["std"] == pytest.approx(0.1) and d["argmax"] == "d1"
    assert res["winners"] == {"a1": 1, "a2": 1}
    a = res["auto_vs_best"]  # 1.5 and 1.0
    assert a["mean"] == pytest.approx(1.25) and a["above_1"] == 1
    assert res["vendor_over_best"]["max"] == pytest.approx(3.0)


def test_analyse_ties_and_missing():
    t = {k: dict(v) for k, v in TABLE.items()}
    t["d2"]["a1"] = 2.5 * 3.0 / 2.0 * 1.2 / 1.2  # a1 ratios 1, 1.25 -> gmean 1.118, within 0.03 of a2
    res = sweep.analyse(_rows(t), epsilon=0.05)["N64_f32"]["families"]["ebpr"]
    assert res["static_ties"] == ["a1", "a2"]
    # a NaN / missing time drops that action from the family's comparison
    rows = _rows(TABLE)
    for r in rows:
        if r["dataset"] == "d2" and r["action"] == "a1":
            r["ms"] = float("nan")
    res = sweep.analyse(rows)["N64_f32"]["families"]["ebpr"]
    assert res["datasets"] == 2 and res["gain_vs_plain"]["n"] == 2  # a1 leaves, the matrices stay
    assert res["gain_vs_plain"]["max"] == pytest.approx(4.0 / 1.2) and "a1" not in res["static_ties"]


def test_render_markdown_lists_every_family_row():
    md = sweep.render_markdown(sweep.analyse(_rows(TABLE)))
    assert "plain / best" in md and "hipSPARSE / best" in md and "best static ebpr: `a2`" in md


def test_actions_are_matrix_independent():
    acts = sweep.spmm_actions(64)
    names = [a.name for a in acts]
    assert len(names) == len(set(names))
    assert {a.family for a in acts} == {"ebpr", "rbpr", "ebsr", "rbsr", "vendor", "select"}
    assert names == [a.name for a in sweep.spmm_actions(64)]
    eb = [a for a in acts if a.family == "ebsr" and a.kind == "cg"][0]
    p = dict(eb.params)
    k = eb.knobs_for({"rows": 262144, "nnz": 4194304})
    g = 1 << p["group_factor"]
    blocks = min(-(-(-(-4194304 // (p["thread_nz"] * g))) // (p["thread_per_block"] // g)),
                 256 * p["waves_per_cu"] * 64 // p["thread_per_block"])
    assert k["block_denom"] == 262144 // blocks and "waves_per_cu" not in k
    assert sweep.spmm_actions(64, families=["ebpr"], vendor=False, select=False)[0].name == "plain_ebpr"


TRACE_HEADER = ('"Kind","Agent_Id","Queue_Id","Stream_Id","Thread_Id","Dispatch_Id","Kernel_Id","Kernel_Name",'
                '"Correlation_Id","Start_Timestamp","End_Timestamp","LDS_Block_Size","Scratch_Size","VGPR_Count",'
                '"Accum_VGPR_Count","SGPR_Count","Workgroup_Size_X","Workgroup_Size_Y","Workgroup_Size_Z",'
                '"Grid_Size_X","Grid_Size_Y","Grid_Size_Z"')
PMC_HEADER = ('"Correlation_Id","Dispatch_Id","Agent_Id","Queue_Id","Process_Id","Thread_Id","Grid_Size","Kernel_Id",'
              '"Kernel_Name","Workgroup_Size","LDS_Block_Size","Scratch_Size","VGPR_Count","Accum_VGPR_Count",'
              '"SGPR_Count","Counter_Name","Counter_Value","Start_Timestamp","End_Timestamp"')


def test_clean_joins_dispatches_to_windows(tmp_path):
    off = 10**9  # the boottime clock is offset: only the monotonic windows contain the dispatches
    man = [{"dataset": "m", "N": 8, "dtype": "f32", "action": a, "family": "ebpr", "kind": "cg",
            "t0": {"monotonic": t0, "boottime": t0 + off}, "t1": {"monotonic": t1, "boottime": t1 + off}}
           for a, t0, t1 in (("x", 100_000, 200_000), ("y", 300_000, 400_000), ("z", 500_000, 600_000))]
    disp = [(1, "before", 50_000, 60_000), (2, "fill", 110_000, 150_000), (3, "sr_kernel", 150_000, 195_000),
            (4, "pr_kernel", 310_000, 390_000), (5, "rb_kernel", 520_000, 540_000)]
    tr = tmp_path / "run_kernel_trace.csv"
    tr.write_text(TRACE_HEADER + "\n" + "".join(
        f'"KERNEL_DISPATCH","Agent 1",1,0,7,{i},9,"{n}",{i},{s},{e},0,0,8,0,32,256,1,1,1024,1,1\n'
        for i, n, s, e in disp))
    pm = tmp_path / "run_counter_collection.csv"
    pm.write_text(PMC_HEADER + "\n" + "".join(
        f'{i},{i},"Agent 1",1,7,7,1024,9,"{n}",256,0,0,8,0,32,"FETCH_SIZE",{10.0 * i},{s},{e}\n'
        for i, n, s, e in disp))
    rows = sweep.clean(man, sweep.dispatches_from_rocprof([tr, pm]))
    assert [r["kernels"] for r in rows] == [2, 1, 1]
    assert rows[0]["gpu_us"] == pytest.approx(85.0) and rows[0]["main_kernel"] == "sr_kernel"
    assert rows[0]["main_us"] == pytest.approx(45.0) and rows[0]["pmc_FETCH_SIZE"] == pytest.approx(50.0)
    assert rows[1]["pmc_FETCH_SIZE"] == pytest.approx(40.0) and rows[2]["main_kernel"] == "rb_kernel"
    assert all(r["clock"] == "monotonic" for r in rows) and rows[2]["family"] == "ebpr"


@pytest.mark.gpu
def test_tiny_sweep_runs_and_checks(gpu, tmp_path):
    out = tmp_path / "s.jsonl"
    rc = sweep.main(["run", "--suite", "tiny", "--N", "32", "--reps", "2", "--check", "--out", str(out)])
    rows = [json.loads(x) for x in out.read_text().splitlines()]
    assert rc == 0
    timed = [r for r in rows if np.isfinite(r["ms"])]
    assert len(timed) >= 0.9 * len(rows)
    assert all(r["ok"] for r in timed)
    res = sweep.analyse(rows)["N32_f32"]
    assert set(res["families"]) == set(sweep.FAMILIES)


def _bucket_rows():
    times = {"d1": {"a1": 1.0, "a2": 2.0}, "d2": {"a1": 1.0, "a2": 1.5}, "d3": {"a1": 2.5, "a2": 1.0},
             "d4": {"a1": 1.0, "a2": 1.2}}
    params = {"a1": {"group_factor": 4, "thread_per_block": 256, "coarsen_factor": 4, "thread_nz": 8,
                     "waves_per_cu": 16, "write_policy": 2},
              "a2": {"group_factor": 3, "tile_factor": 3, "thread_per_block": 256, "block_numer": 1,
                     "block_denom": 2}}
    var = {"a1": "ebsr", "a2": "rbpr"}
    out = []
    for d, t in times.items():  # all in one bucket: even rows (max 8, mean 4), small, N 64
        for a, v in t.items():
            out.append({"dataset": d, "rows": 1000, "cols": 1000, "nnz": 4000, "max_row": 8, "N": 64,
                        "dtype": "f32", "action": a, "family": var[a], "kind": "cg", "variant": var[a],
                        "params": params[a], "ms": v})
    return out


def test_fit_table_and_leave_one_out():
    from segscan_amd import autotune
    rows = _bucket_rows()
    table = autotune.fit_table(rows)
    key = autotune.tune_prefix() + "|spmm|f32|N64|skew0|s"
    assert list(table) == [key]
    e = table[key]  # gmeans: a1 2.5^(1/4) = 1.257, a2 3.6^(1/4) = 1.377
    assert e["action"] == "a1" and e["variant"] == "ebsr" and e["knobs"]["waves_per_cu"] == 16
    assert e["gmean_vs_best"] == pytest.approx(2.5 ** (1 / 4), abs=1e-4)
    assert e["matrices"] == ["d1", "d2", "d3", "d4"]
    # without d1: a1 2.5^(1/3) = 1.357, a2 1.8^(1/3) = 1.216
    assert autotune.fit_table(rows, exclude="d1")[key]["action"] == "a2"
    # leave-one-out picks: d1 -> a2 (2.0), d2 -> a2 (1.5; 2.4^(1/3) = 1.339 < 1.357),
    # d3 -> a1 (2.5), d4 -> a1 (1.0)
    loo = sweep.analyse(rows)["N64_f32"]["table_loo_vs_best"]
    assert loo["mean"] == pytest.approx(1.75) and loo["max"] == pytest.approx(2.5) and loo["argmax"] == "d3"
    # an entry must beat the built-in "auto" by the margin, and rest on >= 3 matrices
    with_auto = list(rows)
    for d, t in (("d1", 1.0), ("d2", 1.0), ("d3", 1.1), ("d4", 1.0)):  # auto gmean 1.1^(1/4) = 1.024 < 1.257
        with_auto.append({**[r for r in rows if r["dataset"] == d][0], "action": "auto", "family": "select",
                          "kind": "auto", "ms": t})
    assert autotune.fit_table(with_auto) == {}
    assert autotune.fit_table(rows, exclude="d1", min_matrices=4) == {}
    assert autotune.fit_table([r for r in rows if r["dataset"] in ("d1", "d2")]) == {}  # two matrices: no entry


def test_lookup_prefers_matrix_entry_and_resolves_grid():
    import torch
    from segscan_amd import autotune
    stats = {"rows": 262144, "nnz": 4194304, "max": 20000, "mean": 16.0}
    assert autotune.skew_class(stats) == 3 and autotune.size_class(stats) == "l"
    assert autotune.skew_class({"rows": 10, "nnz": 40, "max": 36, "mean": 4.0}) == 0  # 36 <= 4*4+32
    assert autotune.skew_class({"rows": 10, "nnz": 40, "max": 60, "mean": 4.0}) == 1  # log2(15) // 4 = 0
    bucket = autotune.table_key("spmm", torch.float32, 64, stats)
    table = {bucket: {"variant": "ebsr", "knobs": {"group_factor": 4, "thread_per_block": 256, "thread_nz": 16,
                                                   "coarsen_factor": 4, "waves_per_cu": 16}}}
    v, k, key = autotune.lookup(table, "spmm", torch.float32, 64, stats, 262144)
    # 4194304 / 256-nnz chunks = 16384 chunks, 16 groups per workgroup -> 1024, capped at 256*16*64/256 = 1024
    assert key == bucket and k["block_denom"] == 262144 // 1024 and "waves_per_cu" not in k
    table[autotune.matrix_key("spmm", torch.float32, 64, stats, 262144)] = {"variant": "rbsr", "knobs": None}
    assert autotune.lookup(table, "spmm", torch.float32, 64, stats, 262144)[0] == "rbsr"
    assert autotune.lookup(table, "spmm", torch.float32, 64, stats)[0] == "ebsr"  # no ncol: bucket only
    assert autotune.lookup(table, "spmm", torch.bfloat16, 64, stats) is None
