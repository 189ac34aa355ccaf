"""The batch plan (segscan_batch_plan_*): forward SpMM, per-graph sum pooling and the
backward for dH = dP[graph] in one launch, against the oracle.

* integer operands (values, X and dP in {-2..2}): H, P, dX and dvals equal the oracle
  bit for bit, on batches with one-node graphs, graphs without edges, rows longer than a
  work item (partials of one row spread over all groups of a workgroup) and items with
  fewer entries than groups, over lane shapes, item sizes and storage types;
* U(-1, 1) operands: normalized error <= 1e-6 against float64;
* hipGraph replays with eager work between them, forward-only runs, optional outputs,
  new values through set_values, and the block-diagonal check at plan creation.
"""
import numpy as np
import pytest
import torch

from segscan_amd import _abi, batch, ops, synth

pytestmark = pytest.mark.gpu


def _edge_batch(seed, int_vals=True):
    """Graphs of 1..3000 nodes, some without edges, one with a 2500-entry row."""
    rng = np.random.default_rng(seed)
    sizes = [1, 64, 3000, 1, 17, 500, 2, 900, 1, 40]
    offsets = np.concatenate(([0], np.cumsum(sizes))).astype(np.int64)
    rps, cols = [np.zeros(1, np.int64)], []
    total = 0
    for g, n in enumerate(sizes):
        if g in (3, 6):
            deg = np.zeros(n, np.int64)  # no edges
        else:
            deg = np.clip(rng.zipf(2.0, size=n), 1, n).astype(np.int64)
            if n == 3000:
                deg[1234] = 2500
                deg[10:40] = 0  # a run of empty rows
        sub = synth.csr_from_degrees(rng, deg, n, values="ones")
        rps.append(sub.rowptr[1:].astype(np.int64) + total)
        cols.append(sub.colind.astype(np.int64) + offsets[g])
        total += sub.nnz
    rowptr = np.concatenate(rps).astype(np.int32)
    col = np.concatenate(cols).astype(np.int32)
    if int_vals:
        vals = rng.integers(-2, 3, size=col.size).astype(np.float32)
    else:
        vals = rng.uniform(-1, 1, size=col.size).astype(np.float32)
    M = int(offsets[-1])
    return synth.HostCsr(M, M, rowptr, col, vals), offsets


def _reference_exact(oracle, csr, offsets, X, dP):
    gi = np.repeat(np.arange(offsets.size - 1), np.diff(offsets)).astype(np.int32)
    H = oracle.spmm_f32(csr.rowptr, csr.colind, csr.vals, X)
    P = oracle.segment_coo_f32(H, gi, offsets.size - 1)
    dHt = dP[gi]
    colptr, rowind, perm = oracle.transpose(csr.nrow, csr.ncol, csr.rowptr, csr.colind)
    dX = oracle.spmm_f32(colptr, rowind, csr.vals[perm], dHt)
    dv64, _ = oracle.sddmm_f64(csr.rowptr, csr.colind, dHt, X)
    return H, P, dX, dv64.astype(np.float32)


def _outputs(csr, n_graphs, N, dev):
    return (torch.full((csr.nrow, N), float("nan"), device=dev), torch.full((n_graphs, N), float("nan"), device=dev),
            torch.full((csr.ncol, N), float("nan"), device=dev),
            torch.full((max(csr.nnz, 1),), float("nan"), device=dev)[: csr.nnz])


PLAN_CASES = [
    (torch.float32, 64, None),
    (torch.bfloat16, 64, None),
    (torch.float16, 64, None),
    (torch.float32, 64, dict(thread_nz=1)),                    # tiny items: heads and tails everywhere
    (torch.bfloat16, 64, dict(group_factor=4, coarsen_factor=4, thread_nz=2)),
    (torch.float32, 3, None),                                  # lanes past N
    (torch.float32, 16, dict(group_factor=3, coarsen_factor=2)),
    (torch.bfloat16, 100, None),                               # 64 lanes x 2 columns, partial last lane
    (torch.float32, 128, dict(block_numer=1, block_denom=16)),  # few workgroups: many items each
    (torch.float16, 33, dict(coarsen_factor=1)),
    (torch.float32, 64, dict(block_numer=1, block_denom=1024)),  # a grid under the XCD count: items in order
]


@pytest.mark.parametrize("case", PLAN_CASES, ids=lambda c: f"{str(c[0]).split('.')[-1]}-N{c[1]}-{c[2]}")
def test_batch_plan_integer_exact(gpu, oracle, case):
    dtype, N, knobs = case
    csr, offsets = _edge_batch(3)
    gb = batch.make_batch(csr, offsets, gpu)
    plan = batch.BatchPlan(gb, N, dtype, knobs)
    info = plan.describe()
    print(f"plan {dtype} N={N} {knobs}: {info}", flush=True)
    assert info["nitems"] >= 1 and info["grid"] >= 1
    rng = np.random.default_rng(N)
    X = rng.integers(-2, 3, size=(csr.nrow, N)).astype(np.float32)
    dP = rng.integers(-2, 3, size=(gb.n_graphs, N)).astype(np.float32)
    H, P, dX, dv = _outputs(csr, gb.n_graphs, N, gpu)
    plan.run(torch.from_numpy(X).to(gpu).to(dtype), torch.from_numpy(dP).to(gpu), P, H=H, dX=dX, dvals=dv)
    torch.cuda.synchronize()
    Hr, Pr, dXr, dvr = _reference_exact(oracle, csr, offsets, X, dP)
    assert np.array_equal(H.cpu().numpy(), Hr)
    assert np.array_equal(P.cpu().numpy(), Pr)
    assert np.array_equal(dX.cpu().numpy(), dXr)
    assert np.array_equal(dv.cpu().numpy(), dvr)


def test_batch_plan_float_replays_and_options(gpu, oracle):
    """U(-1, 1) operands within 1e-6 normalized; graph replays with eager work between them;
    forward-only and optional outputs; new values through set_values."""
    csr, offsets = _edge_batch(4, int_vals=False)
    gb = batch.make_batch(csr, offsets, gpu)
    N = 64
    rng = np.random.default_rng(5)
    plan = batch.BatchPlan(gb, N, torch.float32)
    X = torch.empty((csr.nrow, N), device=gpu)
    dP = torch.empty((gb.n_graphs, N), device=gpu)
    H, P, dX, dv = _outputs(csr, gb.n_graphs, N, gpu)
    gi = np.repeat(np.arange(gb.n_graphs), np.diff(offsets))

    def check(vals, Xh, dPh, want=("H", "P", "dX", "dv")):
        H64, HD = oracle.spmm_f64(csr.rowptr, csr.colind, vals, Xh)
        P64 = np.zeros((gb.n_graphs, N))
        PD = np.zeros_like(P64)
        np.add.at(P64, gi, H64)
        np.add.at(PD, gi, HD)
        colptr, rowind, perm = oracle.transpose(csr.nrow, csr.ncol, csr.rowptr, csr.colind)
        dX64, dXD = oracle.spmm_f64(colptr, rowind, vals[perm], dPh[gi])
        dv64, dvD = oracle.sddmm_f64(csr.rowptr, csr.colind, dPh[gi], Xh)
        got = {"H": (H, H64, HD), "P": (P, P64, PD), "dX": (dX, dX64, dXD), "dv": (dv, dv64, dvD)}
        for k in want:
            t, r, d = got[k]
            assert oracle.normalized_error(t.cpu().numpy(), r, d) <= 1e-6, k

    def new_inputs():
        Xh = rng.uniform(-1, 1, size=(csr.nrow, N)).astype(np.float32)
        dPh = rng.standard_normal((gb.n_graphs, N)).astype(np.float32)
        X.copy_(torch.from_numpy(Xh))
        dP.copy_(torch.from_numpy(dPh))
        return Xh, dPh

    Xh, dPh = new_inputs()
    plan.run(X, dP, P, H=H, dX=dX, dvals=dv)
    torch.cuda.synchronize()
    check(csr.vals, Xh, dPh)
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        plan.run(X, dP, P, H=H, dX=dX, dvals=dv)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        plan.run(X, dP, P, H=H, dX=dX, dvals=dv)
    for rep in range(3):
        Xh, dPh = new_inputs()
        for t in (H, P, dX, dv):
            t.fill_(7.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        check(csr.vals, Xh, dPh)
    # forward only: P (and H) without dP; then P alone
    P.fill_(3.0)
    H.fill_(3.0)
    plan.run(X, None, P, H=H)
    torch.cuda.synchronize()
    check(csr.vals, Xh, dPh, want=("H", "P"))
    P.fill_(3.0)
    plan.run(X, None, P)
    torch.cuda.synchronize()
    check(csr.vals, Xh, dPh, want=("P",))
    # backward without dvals, then new values
    plan.run(X, dP, P, dX=dX)
    torch.cuda.synchronize()
    check(csr.vals, Xh, dPh, want=("P", "dX"))
    new_vals = rng.uniform(-1, 1, size=csr.nnz).astype(np.float32)
    plan.set_values(torch.from_numpy(new_vals).to(gpu))
    plan.run(X, dP, P, H=H, dX=dX, dvals=dv)
    torch.cuda.synchronize()
    check(new_vals, Xh, dPh)


def test_batch_plan_rejects_off_block_entries(gpu):
    csr, offsets = _edge_batch(6)
    cols = csr.colind.copy()
    cols[-1] = 0  # last graph's last entry points into graph 0
    bad = synth.HostCsr(csr.nrow, csr.ncol, csr.rowptr, cols, csr.vals)
    gb = batch.make_batch(bad, offsets, gpu)
    with pytest.raises(_abi.SegscanError) as exc:
        batch.BatchPlan(gb, 16, torch.float32)
    assert exc.value.status == _abi.EINVAL and "block diagonal" in str(exc.value)


def test_batched_step_single_launch_matches_kernel_by_kernel(gpu, oracle):
    """BatchedGcnStep's default (the plan) and its kernel-by-kernel form give the same
    results on a batch of synthetic graphs (normalized error against float64)."""
    csr, offsets = synth.batched_graphs(n_graphs=32, min_nodes=16, max_nodes=5000, max_deg=800, seed=8)
    gb = batch.make_batch(csr, offsets, gpu)
    N = 64
    rng = np.random.default_rng(9)
    X = torch.from_numpy(rng.uniform(-1, 1, size=(csr.nrow, N)).astype(np.float32)).to(gpu).to(torch.bfloat16)
    dP = torch.from_numpy(rng.standard_normal((gb.n_graphs, N)).astype(np.float32)).to(gpu)
    one = batch.BatchedGcnStep(gb, N, dtype=torch.bfloat16)
    many = batch.BatchedGcnStep(gb, N, dtype=torch.bfloat16, single_launch=False)
    one.run(X, dP)
    many.run(X, dP)
    torch.cuda.synchronize()
    for a, b in ((one.H, many.H), (one.P, many.P), (one.dX, many.dX), (one.dvals, many.dvals)):
        a, b = a.cpu().numpy(), b.cpu().numpy()
        assert np.max(np.abs(a - b)) <= 1e-5 * (np.max(np.abs(b)) + 1.0)
