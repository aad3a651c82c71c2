"""Benchmark harness entry points of the CSR SpMV / SpMM path.

``benchmark_spmv_case`` follows benchmarks.py:345-507 of the reference: random
CSR from ``_build_random_csr``, one prepared SpMV timed with
``_benchmark_cuda_op``, checked against torch sparse in the reference's
promoted dtype, and the vendor library timed beside it.  On MI355X the vendor
library is rocSPARSE (``vendor.py``, CSR adaptive algorithm); its numbers fill
the reference's ``cusparse_*`` fields so that scripts reading them keep
working, and ``backend_status["vendor_library"]`` names it.

The SpMM harnesses live beside their operators (``spmm_csr.py``,
``spmm_csr_opt_alg2.py``) as in the reference and are re-exported here.
"""

from __future__ import annotations

import torch

from .. import vendor
from ._common import _benchmark_cuda_op, _build_random_csr, _build_random_dense, _tolerance_for_dtype
from .spmm_csr import benchmark_spmm_case, benchmark_spmm_opt_case, comprehensive_spmm_test
from .spmm_csr_opt_alg2 import benchmark_spmm_opt_alg2_case
from .spmv_csr import (
    SPMV_OP_NON,
    _normalize_spmv_op,
    _spmv_op_to_name,
    _spmv_op_transposes,
    flagsparse_spmv_csr,
    prepare_spmv_csr,
)


def _torch_spmv_reference(data, indices, indptr, x, shape, op_code):
    """torch sparse y = op(A) x in the reference's promoted dtype
    (benchmarks.py:403-426: fp16 / bf16 through fp32)."""
    n_rows = int(shape[0])
    device = data.device
    rows = torch.repeat_interleave(torch.arange(n_rows, device=device, dtype=torch.int64),
                                   (indptr[1:] - indptr[:-1]).to(torch.int64))
    low = data.dtype in (torch.float16, torch.bfloat16)
    vals = data.to(torch.float32) if low else data
    coo = torch.sparse_coo_tensor(torch.stack([rows, indices.to(torch.int64)]), vals, shape,
                                  device=device).coalesce()
    xx = (x.to(torch.float32) if low else x).unsqueeze(1)
    if op_code == SPMV_OP_NON:
        y = torch.sparse.mm(coo, xx)
    else:
        t = coo.t().coalesce()
        if op_code == 2 and vals.is_complex():
            t = torch.sparse_coo_tensor(t.indices(), t.values().conj(), t.shape).coalesce()
        y = torch.sparse.mm(t, xx)
    y = y.squeeze(1)
    return y.to(data.dtype) if low else y


def _vendor_op_matrix(data, indices, indptr, shape, op_code):
    """CSR arrays of op(A) for the vendor call (rocSPARSE is run on op(A) as
    a plain CSR so every algorithm is deterministic)."""
    if op_code == SPMV_OP_NON:
        return data, indices, indptr, shape
    from .spmv_csr import _transpose_csr_for_spmv
    d = data.conj() if (op_code == 2 and data.is_complex()) else data
    return _transpose_csr_for_spmv(d.resolve_conj() if d.is_complex() else d, indices, indptr,
                                   shape)


def benchmark_spmv_case(n_rows=4096, n_cols=4096, nnz=65536, value_dtype=torch.float32,
                        index_dtype=torch.int32, warmup=20, iters=200, block_nnz=256,
                        max_segments=None, run_cusparse=True, transpose=False, op=None,
                        index_fallback_policy="auto"):
    """HIP CSR SpMV vs the vendor library (rocSPARSE) on one random matrix;
    same parameters and result layout as benchmarks.py:345-507."""
    device = torch.device("cuda")
    data, indices, indptr = _build_random_csr(n_rows, n_cols, nnz, value_dtype, index_dtype,
                                              device)
    op_code = _normalize_spmv_op(op, transpose=transpose)
    if op is not None and bool(transpose) and op_code == SPMV_OP_NON:
        raise ValueError("transpose=True conflicts with op=non")
    transpose = _spmv_op_transposes(op_code)
    op_name = _spmv_op_to_name(op_code)
    x_size = n_rows if transpose else n_cols
    y_size = n_cols if transpose else n_rows
    x = _build_random_dense(x_size, value_dtype, device)
    shape = (n_rows, n_cols)
    prepared = prepare_spmv_csr(data, indices, indptr, shape, block_nnz=block_nnz,
                                max_segments=max_segments, op=op_code,
                                index_fallback_policy=index_fallback_policy)
    hip_y, hip_ms = _benchmark_cuda_op(
        lambda: flagsparse_spmv_csr(x=x, prepared=prepared, return_time=False),
        warmup=warmup, iters=iters)
    expected = _torch_spmv_reference(data, indices, indptr, x, shape, op_code)
    atol, rtol = _tolerance_for_dtype(value_dtype)
    hip_match = bool(torch.allclose(hip_y, expected, atol=atol, rtol=rtol))
    hip_max_error = float(torch.max(torch.abs(hip_y - expected)).item()) if y_size > 0 else 0.0

    vendor_ms = vendor_match = vendor_max_error = None
    vendor_reason = None
    if run_cusparse:
        vd, vi, vp, vshape = _vendor_op_matrix(data, indices, indptr, shape, op_code)
        _, run, vendor_reason = vendor.spmv_baseline(vd, vi, vp, vshape, x, lambda y: y.clone())
        if run is not None:
            try:
                values, vendor_ms = _benchmark_cuda_op(run, warmup=warmup, iters=iters)
                vendor_match = bool(torch.allclose(values, expected, atol=atol, rtol=rtol))
                vendor_max_error = (float(torch.max(torch.abs(values - expected)).item())
                                    if y_size > 0 else 0.0)
            except Exception as exc:
                vendor_reason = str(exc)
    speedup = vendor_ms / hip_ms if (vendor_ms is not None and hip_ms > 0) else None
    return {
        "parameters": {
            "n_rows": n_rows, "n_cols": n_cols, "nnz": nnz, "value_dtype": str(value_dtype),
            "index_dtype": str(index_dtype), "op": op_name, "transpose": transpose,
            "index_fallback_policy": str(index_fallback_policy).lower(), "warmup": warmup,
            "iters": iters,
        },
        "performance": {
            "triton_ms": hip_ms,  # the reference's field name; here the HIP kernels
            "hip_ms": hip_ms,
            "cusparse_ms": vendor_ms,
            "triton_speedup_vs_cusparse": speedup,
        },
        "verification": {
            "triton_match_reference": hip_match,
            "triton_max_error": hip_max_error,
            "cusparse_match_reference": vendor_match,
            "cusparse_max_error": vendor_max_error,
        },
        "backend_status": {
            "cusparse_unavailable_reason": vendor_reason,
            "vendor_library": "rocSPARSE (csr adaptive)",
        },
        "samples": {"triton": hip_y, "reference": expected},
    }


__all__ = [
    "benchmark_spmv_case",
    "benchmark_spmm_case",
    "benchmark_spmm_opt_case",
    "benchmark_spmm_opt_alg2_case",
    "comprehensive_spmm_test",
]
