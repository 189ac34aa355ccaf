"""Measurement suite over the five synthetic workloads (one JSON line each).

Lives with the tests because its error columns use the CPU oracle; run it on a
GPU as `python tests/perf_suite.py out.jsonl [legs]` from the repository root.

  poisson      16384 x 16384, nnz/row ~ Poisson(16), N = 16 fp32
  powerlaw     262144 x 262144 Zipf(2.1) ~4.19M nnz, N in {4, 16, 64, 128} fp32, bf16 at 64,
               plus segment_coo sum over the same degrees (E ~4.19M, N = 64)
  banded       120000 x 120000 of 3x3 blocks (~20 nnz/row, bandwidth 600), N in {64, 128},
               B in bf16 / fp16 / fp32
  gcn          D^-1/2 (A + I) D^-1/2, 131072 nodes, degree 20, X 131072 x 128 fp32 / bf16:
               forward, backward (A^T dC, SDDMM), mean-pool segment backward
  batched      256 graphs (64..50000 nodes, Zipf 2.1 degrees clipped at 2000), N = 64 bf16:
               hipGraph replay of forward + pooling + backward (the batch plan's one launch, and
               kernel by kernel) vs the per-graph launch loop (plain, AUTO, table-tuned, and the
               table-tuned loop captured as a hipGraph)

Every timed leg uses HIP events on torch's current stream (median of `reps`
after warm-up) with inputs resident in HBM. Errors are the normalized error
max |C - C64| / (|A| |B|) against the float64 oracle on the same operand values
(the oracle is imported only for these error columns).
"""
from __future__ import annotations

import json
import sys
import time

import numpy as np
import torch

import os
sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from segscan_amd import autotune, batch, ops, synth  # noqa: E402


def _emit(out, rec):
    line = json.dumps(rec)
    print(line, flush=True)
    if out:
        with open(out, "a") as f:
            f.write(line + "\n")


def _time(fn, reps=20, warmup=3):
    """Median device milliseconds of fn() (HIP events; a spin kernel queued before the
    start event hides the host's issue cost, as in the autotuner)."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    return autotune.time_call(fn, reps)


def _b_alg(m, N, s_b):
    k_used = int(np.count_nonzero(np.bincount(m.colind, minlength=m.ncol)))
    return 4 * (m.nrow + 1) + m.nnz * 8 + k_used * N * s_b + m.nrow * N * 4


def _best(A, B, stats, reps=5, limit=None):
    cands = [("auto", None)] + autotune.spmm_candidates(B.size(1), B.dtype, stats)
    if limit:
        cands = cands[:9] + [c for c in cands if c[0] == "ebsr"][:limit]
    res = autotune.tune_spmm(A, B, cands, reps=reps)
    return res[0]["variant"], res[0]["knobs"], res


def _err(oracle, m, C, Bq):
    C64, D = oracle.spmm_f64(m.rowptr, m.colind, m.vals, Bq)
    return oracle.normalized_error(C, C64, D)


def spmm_leg(name, m, N, dtype, dev, oracle, out, reps=20, rng=None, extra=None, plain=True):
    rng = rng or np.random.default_rng(0)
    A = m.to_device(dev)
    B32 = rng.uniform(-1, 1, size=(m.ncol, N)).astype(np.float32)
    B = torch.from_numpy(B32).to(dtype).to(dev)
    stats = autotune.row_stats(m.rowptr)
    v, k, res = _best(A, B, stats)
    C = torch.empty((m.nrow, N), device=dev)
    ws = ops.spmm_workspace(A, N, dtype, v, k)
    ms = _time(lambda: ops.spmm(A, B, v, k, out=C, workspace=ws), reps)
    s_b = 4 if dtype == torch.float32 else 2
    rec = {"leg": name, "nrow": m.nrow, "nnz": m.nnz, "N": N, "dtype": str(dtype).split(".")[-1],
           "tuned": {"variant": v, "knobs": k}, "ms": round(ms, 5),
           "GBps_alg": round(_b_alg(m, N, s_b) / (ms * 1e-3) / 1e9, 1),
           "GFLOPs": round(2 * m.nnz * N / (ms * 1e-3) / 1e9, 1)}
    if plain:
        pl = {}
        for pv in ("plain_ebpr", "plain_rbpr", "plain_ebsr", "plain_rbsr"):
            wsp = ops.spmm_workspace(A, N, dtype, pv)
            pl[pv] = round(_time(lambda: ops.spmm(A, B, pv, out=C, workspace=wsp), max(5, reps // 4)), 5)
        rec["plain_ms"] = pl
    Cq = ops.spmm(A, B, v, k).cpu().numpy()
    Bq = B.float().cpu().numpy()
    rec["err_vs_f64_same_operand"] = _err(oracle, m, Cq, Bq)
    if dtype != torch.float32:
        rec["err_vs_f64_fp32_operand"] = _err(oracle, m, Cq, B32)
    if extra:
        rec.update(extra)
    _emit(out, rec)
    return rec


def run(out=None, legs=("poisson", "powerlaw", "segment", "banded", "gcn", "shard", "batched")):
    from oracle import oracle
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(0)

    if "poisson" in legs:
        m = synth.poisson_csr(seed=0)
        rec = spmm_leg("poisson_16384_N16", m, 16, torch.float32, dev, oracle, out)
        Bh = np.random.default_rng(0).uniform(-1, 1, size=(m.ncol, 16)).astype(np.float32)
        t = time.perf_counter()
        for _ in range(10):
            oracle.spmm_f32(m.rowptr, m.colind, m.vals, Bh)
        _emit(out, {"leg": "poisson_16384_N16_cpu_reference", "ms": round((time.perf_counter() - t) * 100, 3),
                    "cores": 1, "gpu_speedup": round((time.perf_counter() - t) * 100 / rec["ms"], 1)})

    if "powerlaw" in legs or "segment" in legs:
        m = synth.powerlaw_csr(seed=0)
    if "powerlaw" in legs:
        for N in (4, 16, 64, 128):
            spmm_leg(f"powerlaw_262144_N{N}", m, N, torch.float32, dev, oracle, out, plain=(N == 64))
        spmm_leg("powerlaw_262144_N64_bf16", m, 64, torch.bfloat16, dev, oracle, out, plain=False)

    if "segment" in legs:
        deg = m.degrees()
        idx = synth.segment_index_from_degrees(deg)
        E, N = idx.size, 64
        src32 = rng.uniform(-1, 1, size=(E, N)).astype(np.float32)
        for dtype in (torch.float32, torch.bfloat16):
            src = torch.from_numpy(src32).to(dtype).to(dev)
            index = torch.from_numpy(idx).to(dev)
            res = autotune.tune_segment(src, index, deg.size, autotune.segment_candidates(N, dtype, E), reps=5)
            v, k = res[0]["variant"], res[0]["knobs"]
            dst = torch.empty((deg.size, N), device=dev)
            info = ops.segment_describe(E, N, deg.size, dtype, v, k)
            ws = torch.empty(max(1, info["workspace_bytes"]), dtype=torch.uint8, device=dev)
            ms = _time(lambda: ops.segment_coo(src, index, deg.size, v, k, out=dst, workspace=ws))
            s = 4 if dtype == torch.float32 else 2
            b_alg = 4 * E + E * N * s + deg.size * N * 4
            got = ops.segment_coo(src, index, deg.size, v, k).cpu().numpy()
            ref64, absd = oracle.segment_coo_f64(src.float().cpu().numpy(), idx, deg.size)
            plain = {r["variant"]: round(r["ms"], 5) for r in res if r["knobs"] is None}
            _emit(out, {"leg": f"segment_coo_E{E}_N{N}_{str(dtype).split('.')[-1]}", "ms": round(ms, 5),
                        "GBps_alg": round(b_alg / (ms * 1e-3) / 1e9, 1), "tuned": {"variant": v, "knobs": k},
                        "reference_default_ms": plain,
                        "err_vs_f64": oracle.normalized_error(got, ref64, absd)})

    if "banded" in legs:
        mb = synth.banded_block_csr(seed=0)
        for N in (64, 128):
            for dtype in (torch.bfloat16, torch.float16, torch.float32):
                spmm_leg(f"banded_120000_N{N}_{str(dtype).split('.')[-1]}", mb, N, dtype, dev, oracle, out,
                         plain=False)

    if "gcn" in legs:
        mg = synth.gcn_normalized(seed=0)
        from segscan_amd import autograd as ag
        for dtype in (torch.float32, torch.bfloat16):
            A = mg.to_device(dev)
            N = 128
            X = torch.from_numpy(rng.standard_normal((mg.ncol, N)).astype(np.float32)).to(dtype).to(dev)
            dC = torch.from_numpy(rng.standard_normal((mg.nrow, N)).astype(np.float32)).to(dev)
            t0 = time.perf_counter()
            pattern = ag.SpmmPattern(A)
            torch.cuda.synchronize()
            t_pattern = (time.perf_counter() - t0) * 1e3
            vals = A.vals.clone().requires_grad_()
            Xg = X.clone().requires_grad_()

            def fwd_bwd():
                vals.grad = None
                Xg.grad = None
                C = ag.spmm(vals, Xg, pattern)
                C.backward(dC)

            pattern.fused = False
            ms_sep = _time(fwd_bwd, reps=10)  # transposed SpMM + SDDMM as two kernels
            pattern.fused = True
            ms = _time(fwd_bwd, reps=10)
            # the fused backward's shape tuned for this matrix (the built-in one is a candidate)
            Xs = Xg.detach() if dtype != torch.float32 else Xg.detach().float()
            tb = autotune.tune_backward(pattern.T, vals.detach(), dC, Xs, autotune.backward_candidates(N), reps=5)
            pattern.knobs_bwd = tb[0]["knobs"]
            ms_tuned = _time(fwd_bwd, reps=10)
            pattern.knobs_bwd = None
            fwd = _time(lambda: ag.spmm(vals.detach(), Xg.detach(), pattern), reps=10)
            # accuracy of the gradients against float64
            fwd_bwd()
            Xq = Xg.detach().float().cpu().numpy()
            dCh = dC.cpu().numpy().astype(np.float64)
            colptr, rowind, perm = oracle.transpose(mg.nrow, mg.ncol, mg.rowptr, mg.colind)
            dX64 = oracle.spmm_f64d(colptr, rowind, mg.vals.astype(np.float64)[perm], dCh)
            dXabs = oracle.spmm_f64d(colptr, rowind, np.abs(mg.vals.astype(np.float64))[perm], np.abs(dCh))
            dv64 = oracle.sddmm_f64d(mg.rowptr, mg.colind, dCh, Xq.astype(np.float64))
            dvabs = oracle.sddmm_f64d(mg.rowptr, mg.colind, np.abs(dCh), np.abs(Xq.astype(np.float64)))
            # mean-pool segment backward over 512-node groups
            gidx = torch.from_numpy((np.arange(mg.nrow) // 512).astype(np.int32)).to(dev)
            H = torch.from_numpy(rng.standard_normal((mg.nrow, N)).astype(np.float32)).to(dev).requires_grad_()
            ng = (mg.nrow + 511) // 512
            dP = torch.randn((ng, N), device=dev)

            def pool_fb():
                H.grad = None
                ag.segment_pool(H, gidx, ng, mean=True).backward(dP)

            pool_ms = _time(pool_fb, reps=10)
            _emit(out, {"leg": f"gcn_131072_N128_{str(dtype).split('.')[-1]}", "nnz": mg.nnz,
                        "fwd_ms": round(fwd, 4), "fwd_bwd_ms": round(ms, 4), "fwd_bwd_separate_ms": round(ms_sep, 4),
                        "fwd_bwd_tuned_bwd_ms": round(ms_tuned, 4), "bwd_knobs_tuned": tb[0]["knobs"],
                        "transpose_ms_once": round(t_pattern, 3),
                        "meanpool_fwd_bwd_ms": round(pool_ms, 4),
                        "err_dX": oracle.normalized_error(Xg.grad.float().cpu().numpy(), dX64, dXabs),
                        "err_dvals": oracle.normalized_error(vals.grad.cpu().numpy(), dv64, dvabs)})

    if "shard" in legs:
        # One GPU's share of an 8-way row-sharded SpMM: nnz-balanced row blocks of the
        # power-law matrix with full K (B replicated), each timed on its own; the
        # slowest block is what every GPU of an 8-GPU job would wait for.
        m = synth.powerlaw_csr(seed=0)
        N = 64
        B32 = rng.uniform(-1, 1, size=(m.ncol, N)).astype(np.float32)
        cum = m.rowptr.astype(np.int64)
        bounds = [0] + [int(np.searchsorted(cum, k * m.nnz // 8)) for k in range(1, 8)] + [m.nrow]
        for dtype in (torch.float32, torch.bfloat16):
            B = torch.from_numpy(B32).to(dtype).to(dev)
            times, nnzs = [], []
            for k in range(8):
                r0, r1 = bounds[k], bounds[k + 1]
                rp = (m.rowptr[r0:r1 + 1] - m.rowptr[r0]).astype(np.int32)
                sl = slice(int(m.rowptr[r0]), int(m.rowptr[r1]))
                A = ops.CsrMatrix.from_numpy(r1 - r0, m.ncol, rp, m.colind[sl], m.vals[sl], device=dev)
                C = torch.empty((A.nrow, N), device=dev)
                ws = ops.spmm_workspace(A, N, dtype, "auto")
                times.append(_time(lambda: ops.spmm(A, B, "auto", out=C, workspace=ws), 20))
                nnzs.append(A.nnz)
            _emit(out, {"leg": f"shard_1of8_powerlaw_262144_N64_{str(dtype).split('.')[-1]}", "shards": 8,
                        "rows": [bounds[k + 1] - bounds[k] for k in range(8)], "nnz": nnzs,
                        "ms": [round(t, 5) for t in times], "max_ms": round(max(times), 5),
                        "mean_ms": round(float(np.mean(times)), 5)})

    if "batched" in legs:
        csr, offsets = synth.batched_graphs(seed=0)
        b = batch.make_batch(csr, offsets, dev)
        N = 64
        X = torch.from_numpy(rng.uniform(-1, 1, size=(csr.nrow, N)).astype(np.float32)).to(torch.bfloat16).to(dev)
        dP = torch.from_numpy(rng.standard_normal((b.n_graphs, N)).astype(np.float32)).to(dev)
        stats = autotune.row_stats(csr.rowptr)
        v, k, _ = _best(b.A, X, stats, reps=3, limit=60)
        # the batch plan: forward, pooling and backward in one launch
        step = batch.BatchedGcnStep(b, N, dtype=torch.bfloat16, variant=v, knobs=k)
        eager_ms = _time(lambda: step.run(X, dP), reps=10)
        graph_ms = _time(step.capture(X, dP).replay, reps=20)
        # kernel by kernel: SpMM, segment_coo pooling, fused backward reading dP through the graph index
        kbk = batch.BatchedGcnStep(b, N, dtype=torch.bfloat16, variant=v, knobs=k, single_launch=False)
        kbk_graph_ms = _time(kbk.capture(X, dP).replay, reps=20)
        # the fused backward's shape tuned for this batch (the built-in one is a candidate), on
        # dP read through the graph index exactly as the kernel-by-kernel step runs it
        tb = autotune.tune_backward(kbk.T, None, dP, X, autotune.backward_candidates(N), reps=5,
                                    vals_t=kbk.At.vals, dc_map=b.graph_index)
        kbk_t = batch.BatchedGcnStep(b, N, dtype=torch.bfloat16, variant=v, knobs=k, knobs_bwd=tb[0]["knobs"],
                                     single_launch=False)
        kbk_tuned_ms = _time(kbk_t.capture(X, dP).replay, reps=20)
        sep = batch.BatchedGcnStep(b, N, dtype=torch.bfloat16, variant=v, knobs=k, fused=False, single_launch=False)
        graph_sep_ms = _time(sep.capture(X, dP).replay, reps=20)  # transposed SpMM + SDDMM as two kernels
        # forward + pooling: the plan with dP = None against the per-graph launch loops
        fwd_ms = _time(lambda: step.plan.run(X, None, step.P, H=step.H), reps=10)
        fwd_kbk_ms = _time(lambda: (ops.spmm(b.A, X, v, k, out=kbk.H, workspace=kbk.ws_f),
                                    ops.segment_coo(kbk.H, b.graph_index, b.n_graphs, out=kbk.P, workspace=kbk.ws_s)),
                           reps=10)
        H = torch.empty((csr.nrow, N), device=dev)
        P = torch.empty((b.n_graphs, N), device=dev)
        loop = batch.PerGraphLoop(csr, offsets, N, device=dev, variant="plain_ebpr", dtype=torch.bfloat16)
        loop_ms = _time(lambda: loop.run(X, H, P), reps=5, warmup=1)
        loop_auto = batch.PerGraphLoop(csr, offsets, N, device=dev, variant="auto", dtype=torch.bfloat16)
        loop_auto_ms = _time(lambda: loop_auto.run(X, H, P), reps=5, warmup=1)
        # each graph's variant and knobs from the tuning table's bucket (AUTO where none), run
        # eagerly and captured as one hipGraph (no host issue cost left)
        loop_tab = batch.PerGraphLoop(csr, offsets, N, device=dev, tuner=batch.table_tuner(torch.bfloat16, N),
                                      dtype=torch.bfloat16)
        loop_tab_ms = _time(lambda: loop_tab.run(X, H, P), reps=5, warmup=1)
        torch.cuda.synchronize()
        g_loop = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g_loop):
            loop_tab.run(X, H, P)
        loop_tab_graph_ms = _time(g_loop.replay, reps=10)
        _emit(out, {"leg": "batched_256_graphs_N64_bf16", "nodes": csr.nrow, "nnz": csr.nnz,
                    "tuned": {"variant": v, "knobs": k}, "plan": step.plan.describe(),
                    "fwd_bwd_plan_eager_ms": round(eager_ms, 4), "fwd_bwd_plan_hipgraph_ms": round(graph_ms, 4),
                    "fwd_bwd_kernel_by_kernel_hipgraph_ms": round(kbk_graph_ms, 4),
                    "fwd_bwd_kernel_by_kernel_tuned_bwd_ms": round(kbk_tuned_ms, 4),
                    "fwd_bwd_kernel_by_kernel_separate_bwd_ms": round(graph_sep_ms, 4),
                    "bwd_knobs_tuned": tb[0]["knobs"],
                    "bwd_tuning": [{"knobs": r["knobs"], "ms": round(r["ms"], 4)} for r in tb],
                    "fwd_pool_plan_ms": round(fwd_ms, 4), "fwd_pool_kernel_by_kernel_ms": round(fwd_kbk_ms, 4),
                    "fwd_pool_per_graph_loop_plain_ms": round(loop_ms, 4),
                    "fwd_pool_per_graph_loop_auto_ms": round(loop_auto_ms, 4),
                    "fwd_pool_per_graph_loop_table_ms": round(loop_tab_ms, 4),
                    "fwd_pool_per_graph_loop_table_hipgraph_ms": round(loop_tab_graph_ms, 4),
                    "speedup_plan_vs_table_loop_hipgraph": round(loop_tab_graph_ms / fwd_ms, 2)})


if __name__ == "__main__":
    out = sys.argv[1] if len(sys.argv) > 1 else None
    legs = tuple(sys.argv[2].split(",")) if len(sys.argv) > 2 else ("poisson", "powerlaw", "segment", "banded", "gcn",
                                                                    "shard", "batched")
    run(out, legs)
