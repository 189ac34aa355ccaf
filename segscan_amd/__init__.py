"""segscan_amd — CSR SpMM and segment_coo on AMD MI355X (gfx950).

The compute path is libsegscan_hip.so (hand-written HIP kernels behind the C
ABI in include/segscan.h). This package is its Python host layer:

  * ops             torch-facing calls: spmm, segment_coo, csr_transpose,
                    sddmm, segment_gather, segment_counts, device checks
  * reference_api   the reference's launcher names and argument meaning
                    (csrspmm_parreduce_nnzbalance_cg(A, N, B, C, ...), ...)
  * synth           seeded synthetic inputs
  * autotune        knob search per (nnz/row skew, N) bucket with a JSON table
  * autograd        test-layer torch.autograd.Function over the same C ABI
  * batch           block-diagonal batching of many graphs, hipGraph capture

Importing the package does not load the library; the first call does, and it
raises ImportError if the library is not built. There is no CPU fallback.
"""
from ._abi import SegscanError, Knobs, load as load_library  # noqa: F401

__all__ = ["SegscanError", "Knobs", "load_library"]
