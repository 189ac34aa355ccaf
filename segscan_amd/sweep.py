"""Sweep and analysis pipeline for the SpMM variants.

Replaces the reference's measurement scripts with one module and three steps:

  run     time every launch configuration ("action") of every family on every
          matrix of a suite, HIP events on the current stream — the sweep of
          test-cg/src/test.py:42-59 (there under Nsight Compute, one report per
          matrix), with test.py's check mode (:42-48) as ``--check``;
  clean   join rocprofv3 CSV output (kernel trace and/or counter collection) of a
          ``run --profile`` pass to the actions — the role of
          test-cg/src/clean.py:1-49, which flattened .ncu-rep reports to a CSV of
          (metrics, dataset, kernel, action);
  report  the ratio / geometric-mean analyses of test-cg/src/count.py:1-51,
          static.py:1-34 and dgcmp.py:1-30 that produced README tables 4-6:
          gain from tuning, best vs the vendor library, dynamic vs static (best
          static configuration by geometric mean, ties within 0.01), plus the
          winner histogram of count.py:37-40 and the quality of this library's
          own run-time selection ("auto" heuristic and the tuned table).

CLI: ``python -m segscan_amd.sweep run|clean|report ...`` (``--help`` on each).
The matrices are either a synthetic suite or Matrix Market files named in a
comma-separated list (the layout of test.py:29-38: <dir>/<name>/<name>.mtx).
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time
from dataclasses import dataclass
from pathlib import Path
from typing import Callable, Iterable

import numpy as np

FAMILIES = ("ebpr", "rbpr", "ebsr", "rbsr")
VENDOR_ALGS = ("default", "csr_alg1", "csr_alg2", "csr_alg3")


# ---- actions -----------------------------------------------------------------------------------

@dataclass(frozen=True)
class Action:
    """One launch configuration. ``name`` is identical across matrices, so the
    static analysis can compare one configuration over the whole suite; knobs
    that depend on the matrix (grid size) are resolved by ``knobs_for``."""
    name: str
    family: str          # ebpr / rbpr / ebsr / rbsr / vendor / select
    kind: str            # plain / default / cg / vendor / auto / tuned
    variant: str | None = None
    params: tuple = ()   # (key, value) pairs, matrix-independent
    vendor_alg: str | None = None

    def knobs_for(self, stats: dict) -> dict | None:
        if self.kind in ("plain", "default", "vendor", "auto", "tuned"):
            return None
        from .autotune import resolve_knobs
        return resolve_knobs(dict(self.params), stats)


def _label(family: str, p: dict) -> str:
    if family in ("ebpr", "rbpr"):
        t = (p["group_factor"], p["thread_per_block"], p["tile_factor"], p["block_numer"], p["block_denom"])
    elif family == "rbsr":
        t = (p["coarsen_factor"], p["thread_per_block"], p["tile_factor"])
    else:
        t = (p["group_factor"], p["thread_per_block"], p["coarsen_factor"], p["thread_nz"],
             f"w{p['waves_per_cu']}", f"p{p['write_policy']}")
    return f"{family}_cg({','.join(str(x) for x in t)})"


def spmm_actions(N: int, dtype_bytes: int = 4, families: Iterable[str] = FAMILIES, vendor: bool = True,
                 select: bool = True) -> list[Action]:
    """The action list of one (N, dtype): per family the plain kernel, the
    reference default configuration and a knob grid (the grids of
    test-cg/src/params.md mapped to wave64), then the vendor algorithms and the
    library's own selections. The list does not depend on the matrix."""
    vec = 16 // dtype_bytes
    cf = vec
    while cf > 1 and N % cf:
        cf //= 2
    lanes = min(64, -(-N // cf))
    g = 1
    while g < lanes:
        g *= 2
    out: list[Action] = []
    for fam in families:
        out.append(Action(f"plain_{fam}", fam, "plain", f"plain_{fam}"))
        out.append(Action(f"{fam}_default", fam, "default", fam))
        grid: list[dict] = []
        if fam == "ebpr":
            for gf, tf in ((5, 5), (4, 4), (3, 3), (2, 2), (6, 6), (4, 5), (3, 5)):
                for bd in (1, 4, 16):
                    grid.append(dict(group_factor=gf, thread_per_block=256, tile_factor=tf, block_numer=1,
                                     block_denom=bd))
        elif fam == "rbpr":
            for gf, tf in ((3, 3), (2, 2), (4, 4), (5, 5), (6, 6), (2, 3)):
                for bd in (1, 2, 4):
                    grid.append(dict(group_factor=gf, thread_per_block=256, tile_factor=tf, block_numer=1,
                                     block_denom=bd))
        elif fam == "rbsr":
            for c in sorted({1, 2, min(4, cf), cf}):
                for tf in (5, 6):
                    grid.append(dict(coarsen_factor=c, thread_per_block=256, tile_factor=tf))
        else:
            c_, g_ = cf, g
            shapes = []
            while c_ >= 1 and g_ <= 64 and len(shapes) < 3:
                shapes.append((c_, g_))
                c_, g_ = c_ // 2, g_ * 2
            for c_, g_ in shapes:
                gf_ = int(math.log2(g_))
                for tnz in (2, 4, 8, 16, 32):
                    chunk = tnz * g_
                    if chunk < 32 or chunk > 2048:
                        continue
                    for tpb in (256, 512):
                        for waves in (8, 16, 32):
                            grid.append(dict(group_factor=gf_ if gf_ else -1, thread_per_block=tpb,
                                             coarsen_factor=c_, thread_nz=tnz, waves_per_cu=waves,
                                             write_policy=2 if chunk >= 64 else 3))
        seen = set()
        for p in grid:
            name = _label(fam, p)
            if name in seen:
                continue
            seen.add(name)
            out.append(Action(name, fam, "cg", fam, tuple(sorted(p.items()))))
    if vendor:
        out += [Action(f"hipsparse_{a}", "vendor", "vendor", vendor_alg=a) for a in VENDOR_ALGS]
    if select:
        out += [Action("auto", "select", "auto", "auto"), Action("tuned", "select", "tuned")]
    return out


# ---- matrix suites -----------------------------------------------------------------------------

def synthetic_suite(scale: str = "full") -> list[tuple[str, Callable]]:
    """Matrices covering the structure classes of the reference's SuiteSparse set
    (uniform, power-law of several exponents, banded FEM-like blocks, normalized
    graphs, rectangular, many empty rows, one dense row)."""
    from . import synth
    if scale == "tiny":
        return [
            ("poisson_2k_d8", lambda: synth.poisson_csr(2048, 2048, 8.0, seed=1)),
            ("powerlaw_4k_a2.1", lambda: synth.powerlaw_csr(4096, 4096, 2.1, 1000, 40000, seed=2)),
            ("empty_rows_3k", lambda: synth.small_random_csr(3000, 2500, 0.004, 3, empty_rows=0.5)),
        ]

    def dense_row():
        m = synth.poisson_csr(65536, 65536, 8.0, seed=9)
        deg = np.diff(m.rowptr.astype(np.int64))
        deg[1234] = 65536
        return synth.csr_from_degrees(np.random.default_rng(9), deg, 65536)

    return [
        ("poisson_16k_d4", lambda: synth.poisson_csr(16384, 16384, 4.0, seed=1)),
        ("poisson_16k_d64", lambda: synth.poisson_csr(16384, 16384, 64.0, seed=2)),
        ("poisson_128k_d16", lambda: synth.poisson_csr(131072, 131072, 16.0, seed=3)),
        ("poisson_256k_d4", lambda: synth.poisson_csr(262144, 262144, 4.0, seed=4)),
        ("powerlaw_64k_a1.8", lambda: synth.powerlaw_csr(65536, 65536, 1.8, 20000, 1 << 20, seed=5)),
        ("powerlaw_64k_a2.6", lambda: synth.powerlaw_csr(65536, 65536, 2.6, 20000, 1 << 20, seed=6)),
        ("powerlaw_256k_a2.1", lambda: synth.powerlaw_csr(262144, 262144, 2.1, 20000, 1 << 22, seed=0)),
        ("powerlaw_256k_a3.0", lambda: synth.powerlaw_csr(262144, 262144, 3.0, 5000, 1 << 21, seed=7)),
        ("banded_3x3_40k", lambda: synth.banded_block_csr(40000, 7, 100, seed=8)),
        ("gcn_64k", lambda: synth.gcn_normalized(65536, 20.0, 2.1, seed=10)),
        ("rect_256k_x_16k", lambda: synth.poisson_csr(262144, 16384, 8.0, seed=11)),
        ("rect_16k_x_256k", lambda: synth.poisson_csr(16384, 262144, 32.0, seed=12)),
        ("empty_rows_64k", lambda: synth.small_random_csr(65536, 65536, 2e-4, 13, empty_rows=0.6)),
        ("dense_row_64k", dense_row),
    ]


def mtx_suite(directory: str, names: str | None) -> list[tuple[str, Callable]]:
    """Matrix Market files as test.py:29-38 lays them out (<dir>/<name>/<name>.mtx,
    names comma-separated in a file), or every *.mtx under dir when names is None.
    Values are set to one, as the reference's loader does."""
    from .mtx import read_mtx
    d = Path(directory)
    if names:
        lst = sorted(x.strip() for x in Path(names).read_text().replace("\n", ",").split(",") if x.strip())
        paths = [(n, d / n / f"{n}.mtx") for n in lst]
    else:
        paths = [(p.stem, p) for p in sorted(d.rglob("*.mtx"))]
    return [(n, (lambda p=p: read_mtx(p))) for n, p in paths]


# ---- run ---------------------------------------------------------------------------------------

def _clocks() -> dict:
    return {"monotonic": time.clock_gettime_ns(time.CLOCK_MONOTONIC),
            "boottime": time.clock_gettime_ns(time.CLOCK_BOOTTIME)}


def run_sweep(suite: list[tuple[str, Callable]], Ns: Iterable[int], dtype: str = "f32",
              families: Iterable[str] = FAMILIES, reps: int = 5, check: bool = False, profile: bool = False,
              vendor: bool = True, out=None, manifest=None, log=sys.stderr) -> list[dict]:
    """Times every action on every matrix; one row per (matrix, N, action).
    profile=True runs each action once between two device synchronisations and
    writes the host-clock window of each action to ``manifest`` (for ``clean``)."""
    import torch
    from . import autotune, ops
    tdt = {"f32": torch.float32, "bf16": torch.bfloat16, "f16": torch.float16}[dtype]
    dev = torch.device("cuda", torch.cuda.current_device())
    table = autotune.load_table()
    rows = []
    for name, make in suite:
        h = make()
        stats = autotune.row_stats(h.rowptr)
        A = h.to_device(dev)
        for N in Ns:
            g = torch.Generator(device="cpu").manual_seed(N)
            B = (torch.rand((h.ncol, N), generator=g) * 2 - 1).to(dev).to(tdt)
            C = torch.empty((h.nrow, N), dtype=torch.float32, device=dev)
            ref = scale = None
            if check:  # reference product and |A||B| (the error scale) from a plain kernel
                ref = ops.spmm(A, B.float(), variant="plain_rbpr")
                A_abs = A.with_vals(None if A.vals is None else A.vals.abs())
                scale = ops.spmm(A_abs, B.float().abs(), variant="plain_rbpr").clamp_min_(1e-30)
            actions = spmm_actions(N, tdt.itemsize, families, vendor=vendor and dtype == "f32")
            for act in actions:
                row = {"dataset": name, "rows": h.nrow, "cols": h.ncol, "nnz": h.nnz, "N": N, "dtype": dtype,
                       "max_row": stats["max"], "cv": round(stats["cv"], 4), "action": act.name,
                       "family": act.family, "kind": act.kind, "variant": act.variant,
                       "params": dict(act.params) or None}
                try:
                    fn, resolved = _bind(act, A, B, C, stats, table, tdt, N, h.ncol)
                except Exception as exc:  # an unsupported point is recorded, not fatal
                    row.update(ms=float("nan"), error=str(exc).splitlines()[0][:200])
                    rows.append(row)
                    _emit(out, row)
                    continue
                row["resolved"] = resolved
                if profile:
                    torch.cuda.synchronize()
                    t0 = _clocks()
                    fn()
                    torch.cuda.synchronize()
                    t1 = _clocks()
                    if manifest is not None:
                        manifest.write(json.dumps({"dataset": name, "N": N, "dtype": dtype, "action": act.name,
                                                   "family": act.family, "kind": act.kind,
                                                   "t0": t0, "t1": t1}) + "\n")
                        manifest.flush()
                    row["ms"] = float("nan")
                else:
                    row["ms"] = autotune.time_call(fn, reps)
                if check:
                    fn()
                    torch.cuda.synchronize()
                    row["max_err"] = float(((C - ref).abs() / scale).max())
                    row["ok"] = bool(row["max_err"] <= 1e-4)
                rows.append(row)
                _emit(out, row)
            print(f"[sweep] {name} N={N}: {len(actions)} actions", file=log, flush=True)
        del A
        torch.cuda.empty_cache()
    return rows


def _bind(act: Action, A, B, C, stats, table, tdt, N, ncol):
    """A zero-argument callable that performs the action, and the knobs it resolved to."""
    from . import autotune, ops
    if act.kind == "vendor":
        from .vendor import VendorSpmm
        v = VendorSpmm(A, B, C, act.vendor_alg)
        return v, act.vendor_alg
    if act.kind == "tuned":
        hit = autotune.lookup(table, "spmm", tdt, N, stats, ncol)
        variant, knobs = hit[:2] if hit else ("auto", None)
    else:
        variant, knobs = act.variant, act.knobs_for(stats)
    ws = ops.spmm_workspace(A, N, tdt, variant, knobs)
    info = ops.spmm_describe(A, N, tdt, variant, knobs)

    def fn():
        ops.spmm(A, B, variant=variant, knobs=knobs, out=C, workspace=ws)
    return fn, json.dumps({"variant": variant, "knobs": knobs, "group": info["group_size"],
                           "tnz": info["thread_nz"], "grid": info["grid_x"]}, sort_keys=True)


def _emit(out, row):
    if out is not None:
        out.write(json.dumps(row) + "\n")
        out.flush()


# ---- clean -------------------------------------------------------------------------------------

def read_rocprof_csv(path: Path) -> list[dict]:
    import csv
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def dispatches_from_rocprof(paths: Iterable[Path]) -> dict[int, dict]:
    """Dispatch_Id -> {name, start, end, counters{}} from *_kernel_trace.csv and
    *_counter_collection.csv files (rocprofv3 --output-format csv)."""
    out: dict[int, dict] = {}
    for p in paths:
        for r in read_rocprof_csv(Path(p)):
            did = int(r["Dispatch_Id"])
            d = out.setdefault(did, {"name": r["Kernel_Name"], "start": int(r["Start_Timestamp"]),
                                     "end": int(r["End_Timestamp"]), "counters": {}})
            if "Counter_Name" in r and r.get("Counter_Name"):
                d["counters"][r["Counter_Name"]] = d["counters"].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return out


def clean(manifest_rows: list[dict], dispatches: dict[int, dict]) -> list[dict]:
    """Assigns each dispatch to the action whose host window [t0, t1] contains it
    (device synchronisation on both sides of every action makes the windows
    disjoint and complete) and sums per action: kernel count, device time, the
    longest kernel's name and time, and every counter."""
    best = None
    for clock in ("monotonic", "boottime"):
        wins = sorted(((m["t0"][clock], m["t1"][clock], i) for i, m in enumerate(manifest_rows)))
        starts = np.array([w[0] for w in wins], dtype=np.int64)
        assign = {}
        for did, d in dispatches.items():
            k = int(np.searchsorted(starts, d["start"], side="right")) - 1
            if k >= 0 and d["end"] <= wins[k][1]:
                assign[did] = wins[k][2]
        if best is None or len(assign) > len(best[1]):
            best = (clock, assign)
    clock, assign = best
    agg = [{"kernels": 0, "gpu_us": 0.0, "main_kernel": "", "main_us": 0.0, "counters": {}} for _ in manifest_rows]
    for did, i in assign.items():
        d = dispatches[did]
        a = agg[i]
        us = (d["end"] - d["start"]) / 1e3
        a["kernels"] += 1
        a["gpu_us"] += us
        if us > a["main_us"]:
            a["main_us"], a["main_kernel"] = us, d["name"]
        for k, v in d["counters"].items():
            a["counters"][k] = a["counters"].get(k, 0.0) + v
    out = []
    for m, a in zip(manifest_rows, agg):
        row = {k: m[k] for k in ("dataset", "N", "dtype", "action", "family", "kind")}
        row.update(kernels=a["kernels"], gpu_us=round(a["gpu_us"], 3), main_kernel=a["main_kernel"],
                   main_us=round(a["main_us"], 3), clock=clock)
        row.update({f"pmc_{k}": v for k, v in sorted(a["counters"].items())})
        out.append(row)
    return out


# ---- report ------------------------------------------------------------------------------------

def gmean(x) -> float:
    x = np.asarray(x, dtype=np.float64)
    return float(np.exp(np.mean(np.log(x)))) if x.size else float("nan")


def ratio_stats(ratios: dict[str, float]) -> dict:
    """mean / std (population) / geometric mean / max with the dataset of the max,
    and how many ratios exceed 1 (count.py:31-51, dgcmp.py:25-30)."""
    names = list(ratios)
    r = np.array([ratios[n] for n in names], dtype=np.float64)
    if r.size == 0:
        return {"n": 0}
    k = int(np.argmax(r))
    return {"n": int(r.size), "mean": float(r.mean()), "std": float(r.std()), "gmean": gmean(r),
            "max": float(r[k]), "argmax": names[k], "above_1": int((r > 1).sum()),
            "above_1_names": [n for n, v in zip(names, r) if v > 1][:8]}


def _times(rows: list[dict], time_key: str) -> dict:
    """{(N, dtype): {dataset: {action: t}}} for finite times."""
    t: dict = {}
    for r in rows:
        v = r.get(time_key)
        if v is None or not np.isfinite(v) or v <= 0:
            continue
        t.setdefault((r["N"], r["dtype"]), {}).setdefault(r["dataset"], {})[r["action"]] = float(v)
    return t


def analyse(rows: list[dict], time_key: str = "ms", vendor_action: str = "hipsparse_default",
            epsilon: float = 0.01) -> dict:
    """The README 4-6 analyses per (N, dtype) and family; see the module docstring."""
    from .autotune import fit_table, table_key
    fam_of = {r["action"]: (r["family"], r["kind"]) for r in rows}
    ds_stats = {r["dataset"]: {"rows": r["rows"], "nnz": r["nnz"], "max": r["max_row"],
                               "mean": r["nnz"] / max(1, r["rows"])} for r in rows if "rows" in r}
    loo_tables = {d: fit_table(rows, exclude=d) for d in ds_stats} if ds_stats else {}
    result = {}
    for (N, dt), per_ds in sorted(_times(rows, time_key).items()):
        datasets = sorted(per_ds)
        ours = [a for a, (f, k) in fam_of.items() if f in FAMILIES]
        best_all = {d: min(per_ds[d][a] for a in ours if a in per_ds[d]) for d in datasets
                    if any(a in per_ds[d] for a in ours)}
        winners: dict[str, int] = {}
        for d in best_all:
            w = min((per_ds[d][a], a) for a in ours if a in per_ds[d])[1]
            winners[w] = winners.get(w, 0) + 1
        entry = {"datasets": len(datasets), "winners": dict(sorted(winners.items(), key=lambda x: -x[1])),
                 "families": {}}
        for fam in FAMILIES:
            # the family's actions timed on every matrix of this (N, dtype); an action
            # that failed somewhere leaves the comparison rather than the matrix
            present = {a for d in datasets for a in per_ds[d] if fam_of[a][0] == fam}
            acts = sorted(a for a in present if all(a in per_ds[d] for d in datasets))
            ds = datasets if acts else []
            if not ds:
                continue
            best = {d: min(per_ds[d][a] for a in acts) for d in ds}
            fe: dict = {"datasets": len(ds)}
            for key, base in (("gain_vs_plain", f"plain_{fam}"), ("gain_vs_default", f"{fam}_default")):
                if base in acts:
                    fe[key] = ratio_stats({d: per_ds[d][base] / best[d] for d in ds})
            if all(vendor_action in per_ds[d] for d in ds):
                fe["vs_vendor"] = ratio_stats({d: per_ds[d][vendor_action] / best[d] for d in ds})
            g = {a: gmean([per_ds[d][a] / best[d] for d in ds]) for a in acts}
            gmin = min(g.values())
            static = min(g, key=g.get)
            fe["best_static"] = static
            fe["static_ties"] = sorted(a for a, v in g.items() if abs(v - gmin) < epsilon)
            fe["dynamic_vs_static"] = ratio_stats({d: per_ds[d][static] / best[d] for d in ds})
            entry["families"][fam] = fe
        for sel in ("auto", "tuned"):
            ds = [d for d in best_all if sel in per_ds[d]]
            if ds:
                entry[f"{sel}_vs_best"] = ratio_stats({d: per_ds[d][sel] / best_all[d] for d in ds})
        # bucket table fitted without the matrix it is scored on; no bucket entry -> "auto"
        loo = {}
        for d in best_all:
            if d not in loo_tables:
                continue
            e = loo_tables[d].get(table_key("spmm", dt, N, ds_stats[d]))
            t = per_ds[d].get(e["action"]) if e else None
            t = per_ds[d].get("auto") if t is None else t
            if t is not None:
                loo[d] = t / best_all[d]
        if loo:
            entry["table_loo_vs_best"] = ratio_stats(loo)
        vend = [a for a, (f, _) in fam_of.items() if f == "vendor"]
        ds = [d for d in best_all if vendor_action in per_ds[d]]
        if ds:
            entry["vendor_over_best"] = ratio_stats({d: per_ds[d][vendor_action] / best_all[d] for d in ds})
            entry["best_vendor_over_best"] = ratio_stats(
                {d: min(per_ds[d][a] for a in vend if a in per_ds[d]) / best_all[d] for d in ds})
        result[f"N{N}_{dt}"] = entry
    return result


def render_markdown(res: dict) -> str:
    def s(st):
        if not st or st.get("n", 0) == 0:
            return "| - | - | - | - | - |"
        return (f"| {st['mean']:.3f} | {st['std']:.3f} | {st['gmean']:.3f} | {st['max']:.3f} ({st['argmax']}) "
                f"| {st['above_1']}/{st['n']} |")
    lines = []
    for key, e in res.items():
        lines.append(f"## {key} ({e['datasets']} matrices)\n")
        lines.append("| family | ratio | mean | std | geomean | max | >1 |")
        lines.append("| --- | --- | --- | --- | --- | --- | --- |")
        for fam, fe in e["families"].items():
            lines.append(f"| {fam} | plain / best (gain from tuning) " + s(fe.get("gain_vs_plain")))
            lines.append(f"| {fam} | default / best " + s(fe.get("gain_vs_default")))
            if "vs_vendor" in fe:
                lines.append(f"| {fam} | hipSPARSE / best " + s(fe["vs_vendor"]))
            lines.append(f"| {fam} | best static / best (dynamic vs static) " + s(fe["dynamic_vs_static"]))
        for k, label in (("auto_vs_best", "auto heuristic / best"), ("tuned_vs_best", "tuned table / best"),
                         ("table_loo_vs_best", "bucket table, leave-one-out / best"),
                         ("vendor_over_best", "hipSPARSE / best of all"),
                         ("best_vendor_over_best", "best hipSPARSE alg / best of all")):
            if k in e:
                lines.append(f"| all | {label} " + s(e[k]))
        lines.append("")
        for fam, fe in e["families"].items():
            ties = ", ".join(fe["static_ties"][:4])
            lines.append(f"- best static {fam}: `{fe['best_static']}` (within 0.01: {ties})")
        win = ", ".join(f"{a}: {n}" for a, n in list(e["winners"].items())[:8])
        lines.append(f"- winners: {win}\n")
    return "\n".join(lines)


# ---- CLI ---------------------------------------------------------------------------------------

def _read_jsonl(path) -> list[dict]:
    return [json.loads(x) for x in Path(path).read_text().splitlines() if x.strip()]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m segscan_amd.sweep")
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run", help="time every action on every matrix (JSON lines)")
    r.add_argument("--suite", default="synthetic", help="synthetic | tiny | mtx")
    r.add_argument("--mtx-dir")
    r.add_argument("--names", help="comma-separated matrix names (test-cg/src/final_names.txt layout)")
    r.add_argument("--N", default="4,16,64,128")
    r.add_argument("--dtype", default="f32", choices=["f32", "bf16", "f16"])
    r.add_argument("--families", default=",".join(FAMILIES))
    r.add_argument("--reps", type=int, default=5)
    r.add_argument("--check", action="store_true", help="compare every action's C with a plain kernel's")
    r.add_argument("--profile", action="store_true", help="one launch per action, host windows to --manifest")
    r.add_argument("--manifest")
    r.add_argument("--no-vendor", action="store_true")
    r.add_argument("--out", required=True)
    c = sub.add_parser("clean", help="join rocprofv3 CSVs of a --profile run to its actions")
    c.add_argument("--manifest", required=True)
    c.add_argument("--csv", nargs="+", required=True, help="*_kernel_trace.csv / *_counter_collection.csv")
    c.add_argument("--out", required=True)
    p = sub.add_parser("report", help="ratio / geomean tables from run or clean output")
    p.add_argument("table")
    p.add_argument("--time", default="ms", help="time column: ms (run) or gpu_us / main_us (clean)")
    p.add_argument("--vendor", default="hipsparse_default")
    p.add_argument("--json")
    a = ap.parse_args(argv)
    if a.cmd == "run":
        if a.suite == "mtx":
            suite = mtx_suite(a.mtx_dir, a.names)
        else:
            suite = synthetic_suite("tiny" if a.suite == "tiny" else "full")
        Ns = [int(x) for x in a.N.split(",")]
        man = open(a.manifest, "w") if a.manifest else None
        with open(a.out, "w") as out:
            rows = run_sweep(suite, Ns, a.dtype, a.families.split(","), a.reps, a.check, a.profile,
                             not a.no_vendor, out, man)
        if man:
            man.close()
        bad = [r for r in rows if r.get("ok") is False]
        for b in bad:
            print(f"[sweep] CHECK FAILED {b['dataset']} N={b['N']} {b['action']} err={b['max_err']:.3g}",
                  file=sys.stderr)
        return 1 if bad else 0
    if a.cmd == "clean":
        rows = clean(_read_jsonl(a.manifest), dispatches_from_rocprof(a.csv))
        with open(a.out, "w") as out:
            for row in rows:
                out.write(json.dumps(row) + "\n")
        return 0
    rows = _read_jsonl(a.table)
    res = analyse(rows, a.time, a.vendor)
    if a.json:
        Path(a.json).write_text(json.dumps(res, indent=1) + "\n")
    print(render_markdown(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
