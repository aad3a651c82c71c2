"""COO SpMM behind the reference's operator interface (spmm_coo.py).

``flagsparse_spmm_coo(data, row, col, B, shape, block_n=None, block_nnz=256,
out=None, return_time=False)`` keeps the reference's validation order,
messages, launch-config checks and numerics class (spmm_coo.py:312-391,
526-695): fp16 / bf16 computed in fp32, fp32 in fp64, complex64 in
complex128, rounded once to the input dtype.  The reference coalesces
repeated (row, col) pairs, sorts, and runs a row-run Triton kernel; here the
entries are stably sorted by (row, col), a row pointer is formed, and the
gfx950 CSR SpMM runs with the same widened accumulation (repeated pairs are
multiplied separately and summed in the wide type, which equals the
coalesced value up to the last bits of the wide type).
"""

from __future__ import annotations

import time

import torch

from ..sortkey import stable_order
from ._common import SUPPORTED_INDEX_DTYPES, _INDEX_LIMIT_INT32
from .spmm_csr import SUPPORTED_SPMM_VALUE_DTYPES, _select_spmm_alg1_warp_and_factor


def _normalize_spmm_coo_route(route):
    route = "rowrun" if route is None else str(route).lower()
    if route not in ("rowrun", "atomic"):
        raise ValueError("route must be 'rowrun' or 'atomic'")
    return route


def _prepare_spmm_coo_inputs(data, row, col, B, shape):
    """Checks of spmm_coo.py:312-371, in the same order with the same messages."""
    if len(shape) != 2:
        raise ValueError("shape must be a 2-tuple: (n_rows, n_cols)")
    if data.ndim != 1 or row.ndim != 1 or col.ndim != 1:
        raise ValueError("data, row, and col must be 1D tensors")
    if B.ndim != 2:
        raise ValueError("B must be a 2D dense tensor")
    n_rows, n_cols = int(shape[0]), int(shape[1])
    if n_rows < 0 or n_cols < 0:
        raise ValueError("shape dimensions must be non-negative")
    if data.numel() != row.numel() or data.numel() != col.numel():
        raise ValueError("data, row, and col must have the same length (nnz)")
    if B.shape[0] != n_cols:
        raise ValueError(f"B.shape[0] must be n_cols={n_cols}, got {B.shape[0]}")
    if not all(t.is_cuda for t in (data, row, col, B)):
        raise ValueError("data, row, col, and B must be CUDA tensors")
    if not all(t.device == data.device for t in (row, col, B)):
        raise ValueError("data, row, col, and B must be on the same CUDA device")
    if data.dtype not in SUPPORTED_SPMM_VALUE_DTYPES:
        raise TypeError(
            "data dtype must be one of: float16, bfloat16, float32, float64, complex64, complex128")
    if B.dtype != data.dtype:
        raise TypeError("B dtype must match data dtype")
    if row.dtype not in SUPPORTED_INDEX_DTYPES:
        raise TypeError("row dtype must be torch.int32 or torch.int64")
    if col.dtype not in SUPPORTED_INDEX_DTYPES:
        raise TypeError("col dtype must be torch.int32 or torch.int64")
    nnz = data.numel()
    if nnz > _INDEX_LIMIT_INT32:
        raise ValueError("nnz exceeds the int32 range supported by the Triton COO kernel")
    if nnz > 0:
        min_row, max_row = int(row.min().item()), int(row.max().item())
        min_col, max_col = int(col.min().item()), int(col.max().item())
        if min_row < 0 or max_row >= n_rows:
            raise IndexError("row indices out of range for n_rows")
        if min_col < 0 or max_col >= n_cols:
            raise IndexError("col indices out of range for n_cols")
        if max_row > _INDEX_LIMIT_INT32:
            raise ValueError("row indices exceed the int32 range supported by the Triton kernel")
        if max_col > _INDEX_LIMIT_INT32:
            raise ValueError(
                "column indices exceed the int32 range supported by the Triton kernel")
    data, row, col, B = data.contiguous(), row.contiguous(), col.contiguous(), B.contiguous()
    kernel_row = row.to(torch.int32) if row.dtype == torch.int64 else row
    kernel_col = col.to(torch.int32) if col.dtype == torch.int64 else col
    return data, kernel_row, kernel_col, B, n_rows, n_cols, int(B.shape[1])


def _resolve_spmm_coo_launch_config(n_dense_cols, nnz, block_n=None, block_nnz=256):
    """Launch metadata of spmm_coo.py:374-391 (reported; the HIP schedule is
    the CSR SpMM's)."""
    warp_size, factor = _select_spmm_alg1_warp_and_factor(n_dense_cols)
    if block_n is None:
        block_n = warp_size * factor
    if block_nnz is None:
        block_nnz = 256
    if block_n <= 0 or block_nnz <= 0:
        raise ValueError("block_n and block_nnz must be positive when provided")
    return {"block_n": int(block_n), "block_nnz": int(block_nnz),
            "required_nnz_tiles": int(-(-nnz // block_nnz) if nnz > 0 else 0),
            "heuristic_warp_size": int(warp_size), "heuristic_factor": int(factor)}


def _sorted_csr(data, row, col, n_rows):
    """Stable (row, col) order and the row pointer of the sorted entries."""
    if data.numel() == 0:
        return data, col, torch.zeros(n_rows + 1, dtype=torch.int64, device=data.device)
    width = int(col.max().item()) + 1
    key = row.to(torch.int64) * width + col.to(torch.int64)
    order = stable_order(key, n_rows * width)
    counts = torch.bincount(row.to(torch.int64), minlength=n_rows)
    indptr = torch.zeros(n_rows + 1, dtype=torch.int64, device=data.device)
    torch.cumsum(counts, 0, out=indptr[1:])
    return data[order].contiguous(), col[order].contiguous(), indptr


def _run_spmm_coo_route(data, row, col, B, shape, block_n=None, block_nnz=256, out=None,
                        return_time=False, route="rowrun"):
    from .spmm_csr import flagsparse_spmm_csr
    route = _normalize_spmm_coo_route(route)
    if block_n is not None and block_n <= 0:
        raise ValueError("block_n must be positive when provided")
    if block_nnz is not None and block_nnz <= 0:
        raise ValueError("block_nnz must be positive when provided")
    data, krow, kcol, B, n_rows, n_cols, n_dense = _prepare_spmm_coo_inputs(data, row, col, B,
                                                                           shape)
    _resolve_spmm_coo_launch_config(n_dense, data.numel(), block_n=block_n, block_nnz=block_nnz)
    if out is not None:
        if not out.is_cuda:
            raise ValueError("out must be a CUDA tensor")
        if out.device != data.device:
            raise ValueError("out must be on the same CUDA device as the inputs")
        if tuple(out.shape) != (n_rows, n_dense) or out.dtype != data.dtype:
            raise ValueError("out shape/dtype must match result")
    t0 = None
    if return_time:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
    d, c, indptr = _sorted_csr(data, krow, kcol, n_rows)
    C = flagsparse_spmm_csr(d, c, indptr, B, (n_rows, n_cols))
    elapsed_ms = None
    if return_time:
        torch.cuda.synchronize()
        elapsed_ms = (time.perf_counter() - t0) * 1000.0
    if out is not None:
        out.copy_(C)
        C = out
    if return_time:
        return C, elapsed_ms
    return C


def flagsparse_spmm_coo(data, row, col, B, shape, block_n=None, block_nnz=256, out=None,
                        return_time=False):
    """C = A @ B for COO A (spmm_coo.py:671-695)."""
    return _run_spmm_coo_route(data, row, col, B, shape, block_n=block_n, block_nnz=block_nnz,
                               out=out, return_time=return_time, route="rowrun")


__all__ = ["flagsparse_spmm_coo"]
