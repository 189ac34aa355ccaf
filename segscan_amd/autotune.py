"""Autotuner: picks SpMM / segment_coo variant and knobs per (row-length skew,
N, dtype) bucket by timing candidates on the device, and keeps the winners in a
JSON table.

It replaces the reference's brute-force sweeps (test-cg/src/params.md: rb-pr
:1-100, eb_sr :102-177, rb_sr :179-235, eb_pr :237-438, driven under Nsight by
test-cg/src/test.py:42-59) and the static-vs-dynamic analysis
(test-cg/src/static.py:15-35): the search space below contains the
reference's knob points mapped to wave64 plus the gfx950-specific ones (owner
write policy, long chunks, 16-lane groups over a 256-byte row).

Table key: "<arch>|<kind>|<dtype>|N<n_bucket>|skew<skew_bucket>" where
n_bucket = N rounded up to a power of two and skew_bucket = round(log2(max
row length / mean row length)). Timing uses HIP events on torch's current
stream around each candidate (median of `reps` after one warm-up).
"""
from __future__ import annotations

import json
import math
from pathlib import Path
from typing import Iterable

import numpy as np
import torch

from . import ops

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


def skew_bucket(stats: dict) -> int:
    if stats["mean"] <= 0:
        return 0
    return int(round(math.log2(max(1.0, stats["max"] / stats["mean"]))))


def n_bucket(N: int) -> int:
    return 1 << max(0, int(math.ceil(math.log2(max(1, N)))))


def table_key(kind: str, dtype: torch.dtype, N: int, stats: dict, arch: str = "gfx950") -> str:
    dt = {torch.float32: "f32", torch.bfloat16: "bf16", torch.float16: "f16"}[dtype]
    return f"{arch}|{kind}|{dt}|N{n_bucket(N)}|skew{skew_bucket(stats)}"


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
    for c in sorted({1, 2, cf}):
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
        ("segment_coo", None), ("segment_coo_new", None), ("segment_coo_sr", None),
        ("segment_coo_noshmem_sr", None), ("segment_coo_noshmem_lessatom_sr", None),
        ("segment_coo_sorted_sr", None),
    ]
    # short chunks keep all groups streaming inside a narrow window of src (DRAM
    # row locality); long chunks cut the number of chunk-boundary rows
    for tnz in (1, 2, 4, 16, 32, 64, 128):
        chunk = tnz * g
        nchunks = -(-E // chunk)
        for tpb in (256, 512):
            gpb = tpb // g
            for waves_per_cu in (4, 8, 16):
                blocks = max(1, min(-(-nchunks // gpb), (256 * waves_per_cu * 64) // tpb))
                cands.append(("segment_coo_noshmem_lessatom_sr",
                              dict(group_factor=gf if gf else -1, coarsen_factor=cf, thread_nz=tnz,
                                   thread_per_block=tpb, block_numer=1, block_denom=max(1, E // blocks),
                                   write_policy=2 if chunk >= 64 else 3)))
    return cands


# ---- timing ----------------------------------------------------------------------------------

def time_call(fn, reps: int = 10) -> float:
    """Median milliseconds of fn() over reps (after one warm-up), HIP events on the current stream."""
    fn()
    times = []
    for _ in range(reps):
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
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

def load_table(path: Path | str = DEFAULT_TABLE) -> dict:
    p = Path(path)
    if not p.exists():
        return {}
    return json.loads(p.read_text())


def save_table(table: dict, path: Path | str = DEFAULT_TABLE) -> None:
    Path(path).write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")


def lookup(table: dict, key: str) -> tuple[str, dict | None] | None:
    entry = table.get(key)
    if entry is None:
        return None
    return entry["variant"], entry.get("knobs")
