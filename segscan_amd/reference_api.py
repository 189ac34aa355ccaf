"""The reference's launcher interface, by name and argument meaning.

zhang677/segScan exposes header-only templated launchers called directly by its
harnesses; their knobs are template arguments and every output is a device
buffer owned by the caller. The functions below keep those names, argument
order and semantics, with the knobs as keyword-or-positional integers:

  * atomic variants ADD into C / dst, as the reference's kernels do (its
    harness zero-fills the output first, test-cg/include/util/check.cuh:71);
    store variants overwrite it;
  * `nnz` and `N` are taken as given, like the reference's int arguments;
  * errors raise SegscanError (the reference's SpMM launchers check nothing and
    its segment_coo launchers exit() on a CUDA error).

Every call runs the HIP kernels through libsegscan_hip.so on torch's current
stream; the descriptor mirrors SpMatCsrDescr_t (dataloader.hpp:146-168).
"""
from __future__ import annotations

import torch

from . import _abi
from .ops import CsrMatrix, segment_coo as _segment_coo, spmm as _spmm

SpMatCsrDescr_t = CsrMatrix


def _out_view(C: torch.Tensor, rows: int, N: int) -> torch.Tensor:
    """C as a (rows, N) row-major view of the caller's buffer (flat or 2-D)."""
    if C.dtype != torch.float32:
        raise TypeError("output must be float32 (DType = float, check.cuh:12)")
    if C.dim() == 1:
        if C.numel() < rows * N:
            raise ValueError("output buffer too small")
        return C[: rows * N].view(rows, N)
    return C


def _dense_view(B: torch.Tensor, rows: int, N: int) -> torch.Tensor:
    if B.dim() == 1:
        return B[: rows * N].view(rows, N)
    return B


# ---- plain SpMM launchers (test-cg/include/spmm/spmm.cuh) ------------------------------------

def csrspmm_parreduce_nnzbalance(A: CsrMatrix, N: int, B: torch.Tensor, C: torch.Tensor) -> None:
    """spmm.cuh:167-202 (plain ebpr): C += A.B with atomics."""
    _spmm(A, _dense_view(B, A.ncol, N), "plain_ebpr", out=_out_view(C, A.nrow, N), accumulate=True)


def csrspmm_parreduce_rowbalance(A: CsrMatrix, N: int, B: torch.Tensor, C: torch.Tensor) -> None:
    """spmm.cuh:310-343 (plain rbpr): C = A.B (one store per row)."""
    _spmm(A, _dense_view(B, A.ncol, N), "plain_rbpr", out=_out_view(C, A.nrow, N))


def csrspmm_rowcaching_nnzbalance(A: CsrMatrix, N: int, B: torch.Tensor, C: torch.Tensor) -> None:
    """spmm.cuh:522-591 (plain ebsr): C += A.B with atomics."""
    _spmm(A, _dense_view(B, A.ncol, N), "plain_ebsr", out=_out_view(C, A.nrow, N), accumulate=True)


def csrspmm_seqreduce_rowbalance(A: CsrMatrix, N: int, B: torch.Tensor, C: torch.Tensor) -> None:
    """spmm.cuh:621-637 (plain rbsr): C = A.B."""
    _spmm(A, _dense_view(B, A.ncol, N), "plain_rbsr", out=_out_view(C, A.nrow, N))


# ---- tunable SpMM launchers (test-cg/include/spmm/spmm_cg.cuh) --------------------------------

def csrspmm_parreduce_nnzbalance_cg(A: CsrMatrix, N: int, B: torch.Tensor, C: torch.Tensor, group_factor: int,
                                    thread_per_block: int, tile_factor: int, block_numer: int,
                                    block_denom: int) -> None:
    """spmm_cg.cuh:176-218 <group_factor, thread_per_block, tile_factor, block_numer, block_denom>:
    C += A.B with atomics."""
    k = dict(group_factor=group_factor, thread_per_block=thread_per_block, tile_factor=tile_factor,
             block_numer=block_numer, block_denom=block_denom)
    _spmm(A, _dense_view(B, A.ncol, N), "ebpr", knobs=k, out=_out_view(C, A.nrow, N), accumulate=True)


def csrspmm_parreduce_rowbalance_cg(A: CsrMatrix, N: int, B: torch.Tensor, C: torch.Tensor, group_factor: int,
                                    thread_per_block: int, tile_factor: int, block_numer: int,
                                    block_denom: int) -> None:
    """spmm_cg.cuh:333-374: C += A.B (atomicAdd of the group partials, :278)."""
    k = dict(group_factor=group_factor, thread_per_block=thread_per_block, tile_factor=tile_factor,
             block_numer=block_numer, block_denom=block_denom)
    _spmm(A, _dense_view(B, A.ncol, N), "rbpr", knobs=k, out=_out_view(C, A.nrow, N), accumulate=True)


def csrspmm_rowcaching_nnzbalance_cg(A: CsrMatrix, N: int, B: torch.Tensor, C: torch.Tensor, group_factor: int,
                                     thread_per_block: int, CoarsenFactor: int, ThreadNz: int, block_numer: int,
                                     block_denom: int) -> None:
    """spmm_cg.cuh:553-591 <group_factor, thread_per_block, CoarsenFactor, ThreadNz, block_numer,
    block_denom>: C += A.B (atomic flushes, as the reference)."""
    k = dict(group_factor=group_factor, thread_per_block=thread_per_block, coarsen_factor=CoarsenFactor,
             thread_nz=ThreadNz, block_numer=block_numer, block_denom=block_denom,
             write_policy=_abi.WRITE_ATOMIC)
    _spmm(A, _dense_view(B, A.ncol, N), "ebsr", knobs=k, out=_out_view(C, A.nrow, N), accumulate=True)


def csrspmm_seqreduce_rowbalance_cg(A: CsrMatrix, N: int, B: torch.Tensor, C: torch.Tensor, coarsen_factor: int,
                                    thread_per_block: int, tile_factor: int, block_numer: int,
                                    block_denom: int) -> None:
    """spmm_cg.cuh:664-692 <coarsen_factor, thread_per_block, tile_factor, block_numer, block_denom>:
    C = A.B (values read as float; the reference's int truncation, :608, is not reproduced)."""
    k = dict(coarsen_factor=coarsen_factor, thread_per_block=thread_per_block, tile_factor=tile_factor,
             block_numer=block_numer, block_denom=block_denom)
    _spmm(A, _dense_view(B, A.ncol, N), "rbsr", knobs=k, out=_out_view(C, A.nrow, N))


# ---- segment_coo launchers (new-cg/include/segscan/segscan.cuh) -------------------------------

def _seg(variant, src, index, nnz, N, dst, knobs):
    src2 = _dense_view(src, nnz, N)
    dst_len = dst.numel() // N if dst.dim() == 1 else dst.size(0)
    out = _out_view(dst, dst_len, N)
    _segment_coo(src2[:nnz], index[:nnz], dst_len, variant, knobs=knobs, out=out, accumulate=True)


def segment_coo(src, index, nnz: int, N: int, dst, group_factor: int, thread_per_block: int, tile_factor: int,
                block_numer: int, block_denom: int) -> None:
    """segscan.cuh:150-184: dst[index[e]] += src[e] (sorted index), element-parallel tiles."""
    _seg("segment_coo", src, index, nnz, N, dst,
         dict(group_factor=group_factor, thread_per_block=thread_per_block, tile_factor=tile_factor,
              block_numer=block_numer, block_denom=block_denom))


def segment_coo_new(src, index, nnz: int, N: int, dst, group_factor: int, thread_per_block: int, tile_factor: int,
                    blockDimY: int) -> None:
    """segscan.cuh:327-355: elements along x, column panels along y."""
    _seg("segment_coo_new", src, index, nnz, N, dst,
         dict(group_factor=group_factor, thread_per_block=thread_per_block, tile_factor=tile_factor,
              block_dim_y=blockDimY))


def _sr_knobs(group_factor, thread_per_block, CoarsenFactor, ThreadNz, block_numer, block_denom, policy):
    return dict(group_factor=group_factor, thread_per_block=thread_per_block, coarsen_factor=CoarsenFactor,
                thread_nz=ThreadNz, block_numer=block_numer, block_denom=block_denom, write_policy=policy)


def segment_coo_sr(src, index, nnz: int, N: int, dst, group_factor: int, thread_per_block: int,
                   CoarsenFactor: int, ThreadNz: int, block_numer: int, block_denom: int) -> None:
    """segscan.cuh:817-866: staged sequential walk, atomic flushes."""
    _seg("segment_coo_sr", src, index, nnz, N, dst,
         _sr_knobs(group_factor, thread_per_block, CoarsenFactor, ThreadNz, block_numer, block_denom,
                   _abi.WRITE_ATOMIC))


def segment_coo_noshmem_sr(src, index, nnz: int, N: int, dst, group_factor: int, thread_per_block: int,
                           CoarsenFactor: int, ThreadNz: int, block_numer: int, block_denom: int) -> None:
    """segscan.cuh:868-916: register-staged walk, atomic flushes."""
    _seg("segment_coo_noshmem_sr", src, index, nnz, N, dst,
         _sr_knobs(group_factor, thread_per_block, CoarsenFactor, ThreadNz, block_numer, block_denom,
                   _abi.WRITE_ATOMIC))


def segment_coo_noshmem_lessatom_sr(src, index, nnz: int, N: int, dst, group_factor: int, thread_per_block: int,
                                    CoarsenFactor: int, ThreadNz: int, block_numer: int, block_denom: int) -> None:
    """segscan.cuh:918-972: interior segments stored without atomics."""
    _seg("segment_coo_noshmem_lessatom_sr", src, index, nnz, N, dst,
         _sr_knobs(group_factor, thread_per_block, CoarsenFactor, ThreadNz, block_numer, block_denom,
                   _abi.WRITE_LESSATOM))


def segment_coo_sorted_sr(src, index, nnz: int, N: int, dst, CoarsenFactor: int, ThreadNz: int, DtileSize: int,
                          DtileNum: int, BlockDimY: int) -> None:
    """segscan.cuh:974-1008 <CoarsenFactor, ThreadNz, DtileSize, DtileNum, BlockDimY>."""
    _seg("segment_coo_sorted_sr", src, index, nnz, N, dst,
         dict(coarsen_factor=CoarsenFactor, thread_nz=ThreadNz, dtile_size=DtileSize, dtile_num=DtileNum,
              block_dim_y=BlockDimY, write_policy=_abi.WRITE_LESSATOM))
