"""segment_coo parity on the GPU: the six reference variants plus AUTO against
the oracle (new-cg/include/utils/check.cuh:44-52 restated). Integer-valued src
makes every summation order exact, so results must match bit for bit,
including empty segments (exactly 0) and long runs crossing chunk boundaries."""
import numpy as np
import pytest
import torch

from segscan_amd import ops, synth

pytestmark = pytest.mark.gpu

VARIANTS = [
    ("segment_coo", None),
    ("segment_coo_new", None),
    ("segment_coo_sr", None),
    ("segment_coo_noshmem_sr", None),
    ("segment_coo_noshmem_lessatom_sr", None),
    ("segment_coo_sorted_sr", None),
    ("auto", None),
    ("segment_coo", dict(group_factor=6, tile_factor=6)),
    ("segment_coo", dict(group_factor=1, tile_factor=3)),
    ("segment_coo_new", dict(group_factor=4, tile_factor=4, block_dim_y=4)),
    ("segment_coo_sr", dict(group_factor=4, coarsen_factor=4, thread_nz=32, write_policy=2)),
    ("segment_coo_sr", dict(group_factor=6, coarsen_factor=1, thread_nz=2)),
    ("segment_coo_noshmem_lessatom_sr", dict(group_factor=3, coarsen_factor=8, thread_nz=5)),
    ("segment_coo_sorted_sr", dict(coarsen_factor=4, thread_nz=16, dtile_size=16, dtile_num=1, block_dim_y=8)),
    ("segment_coo_sorted_sr", dict(coarsen_factor=1, thread_nz=3, dtile_size=4, dtile_num=2, block_dim_y=3)),
]


def _ids(c):
    v, k = c
    return v + ("" if not k else "-" + "-".join(f"{a[:2]}{b}" for a, b in k.items()))


def _indices():
    rng = np.random.default_rng(7)
    out = []
    # runs from the reference generator shape (check.sh: -r 100 -nnz 5000 -max 100)
    out.append((np.sort(rng.integers(0, 100, size=5000)).astype(np.int32), 100))
    # power-law runs with empty segments in between and at both ends
    deg = np.clip(rng.zipf(2.1, size=400), 1, 3000)
    deg[rng.random(400) < 0.25] = 0
    deg[:3] = 0
    deg[-5:] = 0
    out.append((synth.segment_index_from_degrees(deg), 400))
    # one huge segment
    out.append((np.full(4097, 3, np.int32), 9))
    # a single element
    out.append((np.array([0], np.int32), 1))
    return out


@pytest.mark.parametrize("cfg", VARIANTS, ids=_ids)
def test_segment_integer_exact(gpu, oracle, cfg):
    variant, knobs = cfg
    rng = np.random.default_rng(3)
    for idx, dst_len in _indices():
        for N in (1, 3, 4, 32, 64):
            src = rng.integers(-2, 3, size=(idx.size, N)).astype(np.float32)
            ref = oracle.segment_coo_f32(src, idx, dst_len)
            out = ops.segment_coo(torch.from_numpy(src).to(gpu), torch.from_numpy(idx).to(gpu), dst_len,
                                  variant=variant, knobs=knobs).cpu().numpy()
            assert np.array_equal(out, ref), f"E={idx.size} N={N} max diff {np.abs(out - ref).max()}"


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16, torch.float16])
@pytest.mark.parametrize("variant", ["segment_coo", "segment_coo_sr", "segment_coo_sorted_sr", "auto"])
def test_segment_float(gpu, oracle, dtype, variant):
    rng = np.random.default_rng(4)
    deg = np.clip(rng.zipf(2.1, size=2000), 1, 5000)
    idx = synth.segment_index_from_degrees(deg)
    src32 = rng.uniform(-1, 1, size=(idx.size, 64)).astype(np.float32)
    src = torch.from_numpy(src32).to(dtype)
    rounded = src.float().numpy()
    out = ops.segment_coo(src.to(gpu), torch.from_numpy(idx).to(gpu), deg.size, variant=variant).cpu().numpy()
    ref64, absd = oracle.segment_coo_f64(rounded, idx, deg.size)
    assert oracle.normalized_error(out, ref64, absd) <= 1e-6


@pytest.mark.parametrize("E,N,dst_len", [(600000, 64, 40000), (70000, 300, 5000), (9000, 513, 700), (3, 7, 9)])
def test_segment_auto_stream_shapes(gpu, oracle, E, N, dst_len):
    """"auto" at larger shapes: many chunks with carries, column slices (N beyond 64
    lanes x 4 columns), partial column groups, and fewer rows than one chunk;
    bit-exact on integer values."""
    rng = np.random.default_rng(E)
    idx = np.sort(rng.integers(0, dst_len, size=E)).astype(np.int32)
    src = rng.integers(-2, 3, size=(E, N)).astype(np.float32)
    info = ops.segment_describe(E, N, dst_len, torch.float32, "auto")
    assert info["variant"] == 4 and info["write_policy"] == 2
    ref = oracle.segment_coo_f32(src, idx, dst_len)
    out = ops.segment_coo(torch.from_numpy(src).to(gpu), torch.from_numpy(idx).to(gpu), dst_len,
                          variant="auto").cpu().numpy()
    assert np.array_equal(out, ref)


def test_segment_accumulate(gpu, oracle):
    rng = np.random.default_rng(8)
    idx, dst_len = _indices()[1]
    src = rng.integers(-2, 3, size=(idx.size, 16)).astype(np.float32)
    init = rng.integers(-2, 3, size=(dst_len, 16)).astype(np.float32)
    for variant in ("segment_coo", "segment_coo_sr", "segment_coo_noshmem_lessatom_sr", "auto"):
        out = torch.from_numpy(init.copy()).to(gpu)
        ops.segment_coo(torch.from_numpy(src).to(gpu), torch.from_numpy(idx).to(gpu), dst_len, variant=variant,
                        out=out, accumulate=True)
        ref = init + oracle.segment_coo_f32(src, idx, dst_len)
        assert np.array_equal(out.cpu().numpy(), ref), variant


def test_segment_empty(gpu):
    out = torch.full((5, 4), 7.0, device=gpu)
    ops.segment_coo(torch.zeros((0, 4), device=gpu), torch.zeros(0, dtype=torch.int32, device=gpu), 5, out=out)
    assert torch.count_nonzero(out).item() == 0


def test_check_index_flags(gpu):
    good = torch.tensor([0, 0, 1, 3, 3], dtype=torch.int32, device=gpu)
    assert ops.check_index(good, 4) == 0
    assert ops.check_index(torch.tensor([0, 2, 1], dtype=torch.int32, device=gpu), 4) & 1
    assert ops.check_index(torch.tensor([0, 5], dtype=torch.int32, device=gpu), 4) & 2


@pytest.mark.parametrize("N", [24, 7, 1030])  # 16-byte lanes; one float per lane; columns looped per lane
def test_segment_counts_and_gather(gpu, N):
    rng = np.random.default_rng(9)
    idx, dst_len = _indices()[1]
    counts = ops.segment_counts(torch.from_numpy(idx).to(gpu), dst_len).cpu().numpy()
    assert np.array_equal(counts, np.bincount(idx, minlength=dst_len))
    ddst = rng.standard_normal((dst_len, N)).astype(np.float32)
    g = ops.segment_gather(torch.from_numpy(ddst).to(gpu), torch.from_numpy(idx).to(gpu)).cpu().numpy()
    assert np.array_equal(g, ddst[idx])
    gm = ops.segment_gather(torch.from_numpy(ddst).to(gpu), torch.from_numpy(idx).to(gpu),
                            counts=torch.from_numpy(counts).to(gpu)).cpu().numpy()
    np.testing.assert_allclose(gm, ddst[idx] / np.maximum(counts[idx], 1)[:, None], rtol=1e-6)
