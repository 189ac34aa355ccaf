"""Autotuner: picks SpMM / segment_coo variant and knobs per (N, row-length
skew, size) bucket, and per matrix, and keeps the winners in a JSON table.

It replaces the reference's brute-force sweeps (test-cg/src/params.md: rb-pr
:1-100, eb_sr :102-177, rb_sr :179-235, eb_pr :237-438, driven under Nsight by
test-cg/src/test.py:42-59) and the static-vs-dynamic analysis
(test-cg/src/static.py:15-35): the search space below contains the
reference's knob points mapped to wave64 plus the gfx950-specific ones (owner
write policy, long chunks, 16-lane groups over a 256-byte row).

Two kinds of table entries:
  bucket  "<arch>|rocm<major>.<minor>|<kind>|<dtype>|N<n>|skew<c>|<size>": fitted from a sweep
          (``fit_table``): the configuration with the lowest geometric mean of
          time / per-matrix best over the sweep's matrices in the bucket;
  matrix  bucket key + "|m<rows>x<cols>:<nnz>:<max row>": the winner of a
          per-matrix tuning run (``tune_spmm``), preferred when present.
n = N rounded up to a power of two; c = 0 for even rows (longest row <= 4 x
mean + 32, the rule the C planner uses), else 1 + floor(log2(max / mean) / 4)
capped at 3; size = "s" below 2^20 nonzeros, else "l". Knobs may hold
"waves_per_cu" instead of a block factor; ``resolve_knobs`` turns it into a
grid for the matrix at hand. Timing uses HIP events on torch's current stream
with the host-side issue cost hidden (``time_call``).

Keys start with the library's segscan_tune_prefix() (architecture and ROCm
release); ``load_table`` drops entries of another prefix, so a table tuned under
an older ROCm is re-tuned rather than misapplied. The same JSON files are read
and written by the C ABI's tuning tables (``NativeTable``, segscan_tune_table_*),
whose autotuner (segscan_autotune_spmm) times the same candidate list in C++.
"""
from __future__ import annotations

import ctypes
import json
import math
from pathlib import Path
from typing import Iterable

import numpy as np
import torch

from . import _abi, ops

DEFAULT_TABLE = Path(__file__).resolve().parent / "tuned_gfx950.json"


# ---- buckets -------------------------------------------------------------------------------

def row_stats(rowptr: np.ndarray) -> dict:
    deg = np.diff(np.asarray(rowptr, dtype=np.int64))
    mean = float(deg.mean()) if deg.size else 0.0
    return {
        "rows": int(deg.size),
        "nnz": int(deg.sum()),
        "mean": mean,
        "max": int(deg.max()) if deg.size else 0,
        "cv": float(deg.std() / mean) if mean > 0 else 0.0,
        "empty_frac": float((deg == 0).mean()) if deg.size else 0.0,
    }


def skew_class(stats: dict) -> int:
    mean, mx = stats["mean"], stats["max"]
    if mean <= 0 or mx <= 4 * mean + 32:
        return 0
    return 1 + min(2, int(math.log2(mx / mean) // 4))


def size_class(stats: dict) -> str:
    return "s" if stats["nnz"] < (1 << 20) else "l"


def n_bucket(N: int) -> int:
    return 1 << max(0, int(math.ceil(math.log2(max(1, N)))))


_DT = {torch.float32: "f32", torch.bfloat16: "bf16", torch.float16: "f16"}


def tune_prefix() -> str:
    """Architecture and ROCm release every key starts with, e.g. "gfx950|rocm7.2"."""
    return _abi.load().segscan_tune_prefix().decode()


def table_key(kind: str, dtype, N: int, stats: dict, prefix: str | None = None) -> str:
    dt = _DT.get(dtype, dtype)
    pre = tune_prefix() if prefix is None else prefix
    return f"{pre}|{kind}|{dt}|N{n_bucket(N)}|skew{skew_class(stats)}|{size_class(stats)}"


def matrix_key(kind: str, dtype, N: int, stats: dict, ncol: int, prefix: str | None = None) -> str:
    return table_key(kind, dtype, N, stats, prefix) + f"|m{stats['rows']}x{ncol}:{stats['nnz']}:{stats['max']}"


def resolve_knobs(knobs: dict | None, stats: dict) -> dict | None:
    """Concrete knobs for one matrix: "waves_per_cu" becomes a grid of
    min(chunks / groups per workgroup, 256 CUs x waves x 64 lanes / workgroup)
    workgroups, expressed as the block factor 1 / (rows / grid)."""
    if not knobs or "waves_per_cu" not in knobs:
        return knobs
    p = dict(knobs)
    waves = p.pop("waves_per_cu")
    gf = p.get("group_factor", 0)
    g = 1 << gf if gf > 0 else 1
    chunk = p["thread_nz"] * g
    nchunks = -(-max(1, stats["nnz"]) // chunk)
    gpb = max(1, p["thread_per_block"] // g)
    blocks = max(1, min(-(-nchunks // gpb), (256 * waves * 64) // p["thread_per_block"]))
    p["block_numer"] = 1
    p["block_denom"] = max(1, stats["rows"] // blocks)
    return p


# ---- search spaces -------------------------------------------------------------------------

def _blocks_to_bd(rows: int, blocks: int) -> int:
    return max(1, rows // max(1, blocks))


def spmm_candidates(N: int, dtype: torch.dtype, stats: dict) -> list[tuple[str, dict | None]]:
    """Reference knob points (params.md) mapped to wave64, plus gfx950 points."""
    rows = max(1, stats["rows"])
    nnz = max(1, stats["nnz"])
    vec = 4 if dtype == torch.float32 else 8
    cf = vec
    while cf > 1 and N % cf:
        cf //= 2
    lanes = min(64, -(-N // cf))
    g = 1
    while g < lanes:
        g *= 2
    gf = int(math.log2(g))
    cands: list[tuple[str, dict | None]] = [
        ("plain_ebpr", None), ("plain_rbpr", None), ("plain_ebsr", None), ("plain_rbsr", None),
        ("ebpr", None), ("rbpr", None), ("ebsr", None), ("rbsr", None),
    ]
    # eb_pr / rb_pr: (group, tile) points of params.md, tile up to a full wave
    for gfac, tfac in ((5, 5), (4, 4), (3, 3), (2, 2), (6, 6), (4, 5), (3, 5)):
        cands.append(("ebpr", dict(group_factor=gfac, tile_factor=tfac, thread_per_block=256,
                                   block_numer=1, block_denom=16)))
    for gfac, tfac in ((3, 3), (2, 2), (4, 4), (5, 5), (6, 6), (2, 3)):
        cands.append(("rbpr", dict(group_factor=gfac, tile_factor=tfac, thread_per_block=256,
                                   block_numer=1, block_denom=2)))
    # rb_sr: columns per lane x slice width
    for c in sorted({1, 2, min(4, cf), cf}):
        for tfac in (5, 6):
            cands.append(("rbsr", dict(coarsen_factor=c, tile_factor=tfac, thread_per_block=256)))
    # eb_sr: a group of lanes over one dense row (128-bit loads, or 64/32-bit with
    # twice / four times the lanes), chunk length, workgroup size, grid size and
    # write policy
    shapes = []
    c_, g_ = cf, g
    while c_ >= 1 and g_ <= 64 and len(shapes) < 3:
        shapes.append((c_, g_))
        c_, g_ = c_ // 2, g_ * 2
    for c_, g_ in shapes:
        gf_ = int(math.log2(g_))
        for tnz in (2, 4, 8, 16, 32, 64):
            chunk = tnz * g_
            if chunk < 32 or chunk > 2048:
                continue
            nchunks = -(-nnz // chunk)
            for tpb in (256, 512):
                gpb = tpb // g_
                for waves_per_cu in (4, 8, 16):
                    blocks = max(1, min(-(-nchunks // gpb), (256 * waves_per_cu * 64) // tpb))
                    for policy in ((2, 3) if chunk >= 64 else (3, 1)):
                        cands.append(("ebsr", dict(group_factor=gf_ if gf_ else -1, coarsen_factor=c_,
                                                   thread_nz=tnz, thread_per_block=tpb, block_numer=1,
                                                   block_denom=_blocks_to_bd(rows, blocks), write_policy=policy)))
    return cands


def segment_candidates(N: int, dtype: torch.dtype, E: int) -> list[tuple[str, dict | None]]:
    """The six reference launchers' defaults, the built-in choice, and the
    sequential kernel over group shapes (128-bit loads, and half-width loads over
    twice the lanes, which 16-bit sources prefer), chunk lengths, workgroup sizes
    and grid sizes."""
    vec = 4 if dtype == torch.float32 else 8
    cf = vec
    while cf > 1 and N % cf:
        cf //= 2
    cands: list[tuple[str, dict | None]] = [
        ("segment_coo", None), ("segment_coo_new", None), ("segment_coo_sr", None),
        ("segment_coo_noshmem_sr", None), ("segment_coo_noshmem_lessatom_sr", None),
        ("segment_coo_sorted_sr", None), ("auto", None),
    ]
    shapes = []
    c_ = cf
    while c_ >= 1 and len(shapes) < 2:
        lanes = min(64, -(-N // c_))
        g_ = 1
        while g_ < lanes:
            g_ *= 2
        shapes.append((c_, g_))
        c_ //= 2
    # short chunks keep all groups streaming inside a narrow window of src (DRAM
    # row locality); long chunks cut the number of chunk-boundary rows
    for c_, g in shapes:
        gf = int(math.log2(g))
        for tnz in (1, 2, 4, 16, 32, 64, 128):
            chunk = tnz * g
            nchunks = -(-E // chunk)
            for tpb in (256, 512):
                gpb = tpb // g
                for waves_per_cu in (4, 8, 16):
                    blocks = max(1, min(-(-nchunks // gpb), (256 * waves_per_cu * 64) // tpb))
                    cands.append(("segment_coo_noshmem_lessatom_sr",
                                  dict(group_factor=gf if gf else -1, coarsen_factor=c_, thread_nz=tnz,
                                       thread_per_block=tpb, block_numer=1, block_denom=max(1, E // blocks),
                                       write_policy=2 if chunk >= 64 else 3)))
    return cands


# ---- timing ----------------------------------------------------------------------------------

_SLEEP_CYCLES_PER_MS: dict[int, float] = {}


def _sleep_cycles_per_ms() -> float:
    """Rate of torch.cuda._sleep (a spin kernel) on the current device, measured once."""
    dev = torch.cuda.current_device()
    if dev not in _SLEEP_CYCLES_PER_MS:
        cycles = 1_000_000
        torch.cuda._sleep(cycles)
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        s.record()
        torch.cuda._sleep(cycles)
        e.record()
        e.synchronize()
        _SLEEP_CYCLES_PER_MS[dev] = cycles / max(1e-3, s.elapsed_time(e))
    return _SLEEP_CYCLES_PER_MS[dev]


def time_call(fn, reps: int = 10, hide_host: bool = True) -> float:
    """Median milliseconds of fn() over reps (after one warm-up), HIP events on the
    current stream. With hide_host, a 0.3 ms spin kernel is queued before the start
    event, so the host-side cost of issuing fn() overlaps it and the interval holds
    only the device work (kernels and the gaps between them)."""
    fn()
    cycles = int(0.3 * _sleep_cycles_per_ms()) if hide_host else 0
    times = []
    for _ in range(reps):
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        if cycles:
            torch.cuda._sleep(cycles)
        s.record()
        fn()
        e.record()
        e.synchronize()
        times.append(s.elapsed_time(e))
    return float(np.median(times))


def tune_spmm(A: ops.CsrMatrix, B: torch.Tensor, candidates: Iterable, reps: int = 10) -> list[dict]:
    C = torch.empty((A.nrow, B.size(1)), dtype=torch.float32, device=B.device)
    results = []
    for variant, knobs in candidates:
        try:
            ws = ops.spmm_workspace(A, B.size(1), B.dtype, variant, knobs)
            ms = time_call(lambda: ops.spmm(A, B, variant=variant, knobs=knobs, out=C, workspace=ws), reps)
        except Exception as exc:  # unsupported knob points are skipped, not fatal
            results.append({"variant": variant, "knobs": knobs, "ms": float("inf"), "error": str(exc)})
            continue
        results.append({"variant": variant, "knobs": knobs, "ms": ms})
    results.sort(key=lambda r: r["ms"])
    return results


def backward_candidates(N: int) -> list[dict | None]:
    """Shapes for the fused backward: the built-in one, then 2 / 4 / 8 columns per lane
    (the group covers all N columns, so at most 64 lanes) with 8- or 16-tile chunks.
    The best shape depends on the matrix: for example, 8 columns per lane wins on a
    GCN adjacency at N = 128 but loses on a power-law matrix."""
    cands: list[dict | None] = [None]
    for cf in (2, 4, 8):
        if -(-N // cf) > 64:
            continue
        for tnz in (8, 16):
            cands.append({"coarsen_factor": cf, "thread_nz": tnz})
    return cands


def tune_backward(T: ops.CscMatrix, vals: torch.Tensor | None, dC: torch.Tensor, B: torch.Tensor,
                  candidates: Iterable, reps: int = 10, vals_t: torch.Tensor | None = None,
                  dc_map: torch.Tensor | None = None) -> list[dict]:
    """Times the fused backward (both dB and dvals) over knob points, fastest first;
    points the library rejects are kept with ms = inf and the error. With dc_map, dC
    is read through the row map exactly as the caller will run it (e.g. dC = dP with
    the pooling index), so the shape is tuned for that access pattern."""
    N = B.size(1)
    dB = torch.empty((T.t.nrow, N), dtype=torch.float32, device=B.device)
    dv = torch.empty(max(T.t.nnz, 1), dtype=torch.float32, device=B.device)[: T.t.nnz]
    results = []
    for knobs in candidates:
        try:
            ws = ops.spmm_backward_workspace(T, N, dC.dtype, knobs, B.device)
            ms = time_call(lambda: ops.spmm_backward(T, vals, dC, B, dB=dB, dvals=dv, knobs=knobs, workspace=ws,
                                                     vals_t=vals_t, dc_map=dc_map), reps)
        except _abi.SegscanError as exc:  # unsupported knob points are skipped, not fatal
            results.append({"knobs": knobs, "ms": float("inf"), "error": str(exc)})
            continue
        results.append({"knobs": knobs, "ms": ms})
    results.sort(key=lambda r: r["ms"])
    return results


def tune_segment(src: torch.Tensor, index: torch.Tensor, dst_len: int, candidates: Iterable, reps: int = 10):
    out = torch.empty((dst_len, src.size(1)), dtype=torch.float32, device=src.device)
    results = []
    for variant, knobs in candidates:
        try:
            info = ops.segment_describe(src.size(0), src.size(1), dst_len, src.dtype, variant, knobs)
            ws = torch.empty(max(1, info["workspace_bytes"]), dtype=torch.uint8, device=src.device)
            ms = time_call(lambda: ops.segment_coo(src, index, dst_len, variant=variant, knobs=knobs, out=out,
                                                   workspace=ws), reps)
        except Exception as exc:
            results.append({"variant": variant, "knobs": knobs, "ms": float("inf"), "error": str(exc)})
            continue
        results.append({"variant": variant, "knobs": knobs, "ms": ms})
    results.sort(key=lambda r: r["ms"])
    return results


# ---- table ---------------------------------------------------------------------------------

def load_table(path: Path | str = DEFAULT_TABLE, stale: list | None = None) -> dict:
    """The table's entries for this architecture and ROCm release; keys of any other
    prefix (older releases, other GPUs) are left out and appended to ``stale``."""
    p = Path(path)
    if not p.exists():
        return {}
    pre = tune_prefix() + "|"
    table = {}
    for k, v in json.loads(p.read_text()).items():
        if k.startswith(pre):
            table[k] = v
        elif stale is not None:
            stale.append(k)
    return table


def save_table(table: dict, path: Path | str = DEFAULT_TABLE) -> None:
    Path(path).write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")


def lookup(table: dict, kind: str, dtype, N: int, stats: dict, ncol: int | None = None):
    """(variant, concrete knobs, key) from the matrix entry if present, else the
    bucket entry, else None."""
    keys = ([matrix_key(kind, dtype, N, stats, ncol)] if ncol is not None else []) + [table_key(kind, dtype, N, stats)]
    for key in keys:
        entry = table.get(key)
        if entry is not None:
            return entry["variant"], resolve_knobs(entry.get("knobs"), stats), key
    return None


class NativeTable:
    """The C ABI's tuning table (segscan_tune_table_*): load / save the JSON files this
    module writes, look up a matrix, and run the C++ autotuner (segscan_autotune_spmm)
    on device data, recording the winner under the matrix's key."""

    def __init__(self, path: Path | str | None = None):
        lib = _abi.load()
        h = ctypes.c_void_p()
        _abi.check(lib.segscan_tune_table_create(ctypes.byref(h)), "tune_table_create")
        self._h = h
        self.loaded = self.dropped = 0
        if path is not None:
            self.load(path)

    def load(self, path) -> None:
        nl, nd = ctypes.c_int64(), ctypes.c_int64()
        _abi.check(_abi.load().segscan_tune_table_load(self._h, str(path).encode(), ctypes.byref(nl), ctypes.byref(nd)),
                   "tune_table_load")
        self.loaded, self.dropped = nl.value, nd.value

    def save(self, path) -> None:
        _abi.check(_abi.load().segscan_tune_table_save(self._h, str(path).encode()), "tune_table_save")

    def lookup(self, A: ops.CsrMatrix, N: int, dtype=torch.float32):
        """(variant id, knobs dict, found: 0 none / 1 bucket / 2 matrix)."""
        v, k, f = ctypes.c_int32(), _abi.Knobs(), ctypes.c_int32()
        _abi.check(_abi.load().segscan_tune_table_lookup(self._h, ctypes.byref(A.desc()), N, ops._DTYPES[dtype],
                                                         ctypes.byref(v), ctypes.byref(k), ctypes.byref(f)),
                   "tune_table_lookup")
        return v.value, k.as_dict(), f.value

    def autotune(self, A: ops.CsrMatrix, B: torch.Tensor, reps: int = 7, record: bool = True):
        """Times the candidates on (A, B) in C++; returns (variant id, knobs dict, ms)."""
        lib = _abi.load()
        N = B.size(1)
        desc = A.desc()
        ws_bytes = ctypes.c_size_t()
        _abi.check(lib.segscan_autotune_workspace_size(ctypes.byref(desc), N, ops._dtype_code(B),
                                                       ctypes.byref(ws_bytes)), "autotune_workspace_size")
        ws = torch.empty(max(1, ws_bytes.value), dtype=torch.uint8, device=B.device)
        C = torch.empty((A.nrow, N), dtype=torch.float32, device=B.device)
        v, k, ms = ctypes.c_int32(), _abi.Knobs(), ctypes.c_float()
        _abi.check(lib.segscan_autotune_spmm(self._h if record else None, ctypes.byref(desc), N, ops._ptr(B),
                                             ops._leading_dim(B, "B"), ops._dtype_code(B), ops._ptr(C), N, reps,
                                             ops._ptr(ws), ws.numel(), ops._stream(None), ctypes.byref(v),
                                             ctypes.byref(k), ctypes.byref(ms)), "autotune_spmm")
        return v.value, k.as_dict(), ms.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _abi.load().segscan_tune_table_destroy(h)
            except Exception:
                pass
            self._h = None


def fit_table(rows: list[dict], kind: str = "spmm", exclude: str | None = None, margin: float = 1.03,
              min_matrices: int = 3) -> dict:
    """Bucket entries from sweep rows (segscan_amd.sweep run output): per (dtype,
    N bucket, skew class, size class) the action with the lowest geometric mean of
    time / per-matrix best among the actions timed on every matrix of the bucket,
    kept only where it beats the library's built-in "auto" choice on those
    matrices by more than ``margin`` and fitted on at least ``min_matrices``
    matrices (elsewhere lookup falls back to "auto"). Three: left out of a
    14-matrix sweep, a matrix judged by entries fitted on two others came out
    6-11 % further from its best than under "auto" at N = 64 and 128; with three
    it is within 1 % of "auto" or better.
    ``exclude`` leaves one matrix out (for leave-one-out evaluation)."""
    families = ("ebpr", "rbpr", "ebsr", "rbsr")
    groups: dict = {}
    auto: dict = {}
    spec: dict = {}
    for r in rows:
        if r["dataset"] == exclude or (r.get("family") not in families and r.get("action") != "auto"):
            continue
        t = r.get("ms")
        if t is None or not np.isfinite(t) or t <= 0:
            continue
        stats = {"rows": r["rows"], "nnz": r["nnz"], "max": r["max_row"],
                 "mean": r["nnz"] / max(1, r["rows"])}
        key = table_key(kind, r["dtype"], r["N"], stats)
        if r["action"] == "auto":
            auto.setdefault(key, {})[r["dataset"]] = float(t)
            continue
        groups.setdefault(key, {}).setdefault(r["dataset"], {})[r["action"]] = float(t)
        spec[r["action"]] = r
    table = {}
    for key, per_ds in groups.items():
        common = set.intersection(*(set(v) for v in per_ds.values()))
        if not common or len(per_ds) < min_matrices:
            continue
        best = {d: min(v.values()) for d, v in per_ds.items()}
        score = {a: float(np.exp(np.mean([np.log(per_ds[d][a] / best[d]) for d in per_ds]))) for a in common}
        a = min(sorted(score), key=score.get)
        au = auto.get(key, {})
        if au and all(d in au for d in per_ds):
            auto_score = float(np.exp(np.mean([np.log(au[d] / best[d]) for d in per_ds])))
            if score[a] * margin >= auto_score:
                continue
        r = spec[a]
        res = json.loads(r["resolved"]) if r.get("resolved") else {}
        table[key] = {"variant": res.get("variant", r.get("variant")), "knobs": r.get("params") or res.get("knobs"),
                      "action": a, "gmean_vs_best": round(score[a], 4), "matrices": sorted(per_ds)}
    return table


def _main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser(prog="python -m segscan_amd.autotune")
    sub = ap.add_subparsers(dest="cmd", required=True)
    f = sub.add_parser("fit", help="write bucket entries fitted from sweep output into the table")
    f.add_argument("sweep", nargs="+", help="segscan_amd.sweep run output (JSON lines)")
    f.add_argument("--table", default=str(DEFAULT_TABLE))
    a = ap.parse_args(argv)
    rows = [json.loads(x) for p in a.sweep for x in Path(p).read_text().splitlines() if x.strip()]
    table = {k: v for k, v in load_table(a.table).items() if "|m" in k}  # keep per-matrix entries
    fitted = fit_table(rows)
    table.update(fitted)
    save_table(table, a.table)
    for k, v in sorted(fitted.items()):
        print(f"{k}: {v['action']} (gmean {v['gmean_vs_best']}, {len(v['matrices'])} matrices)")
    return 0


if __name__ == "__main__":
    raise SystemExit(_main())
