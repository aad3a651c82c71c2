"""Shared dtype tables, tolerances and helpers of the CSR operator modules.

Restates the parts of the reference's ``_common.py`` the CSR SpMV / SpMM path
uses (value / index dtype tables :25-36, ``_tolerance_for_dtype`` :123-134,
random builders :203-238, ``_benchmark_cuda_op`` :353-371).  CuPy bridges are
not part of this package (no CuPy on ROCm).
"""

from __future__ import annotations

import time

import torch

_TORCH_COMPLEX32_DTYPE = getattr(torch, "complex32", None) or getattr(torch, "chalf", None)

SUPPORTED_VALUE_DTYPES = tuple(
    [torch.float16, torch.bfloat16, torch.float32, torch.float64]
    + ([_TORCH_COMPLEX32_DTYPE] if _TORCH_COMPLEX32_DTYPE is not None else [])
    + [torch.complex64, torch.complex128]
)
SUPPORTED_INDEX_DTYPES = (torch.int32, torch.int64)
_INDEX_LIMIT_INT32 = 2**31 - 1

# CuPy is not available on ROCm; kept for name compatibility.
cp = None
cpx_sparse = None


def _torch_complex32_dtype():
    return _TORCH_COMPLEX32_DTYPE


def _is_complex32_dtype(value_dtype) -> bool:
    return _TORCH_COMPLEX32_DTYPE is not None and value_dtype == _TORCH_COMPLEX32_DTYPE


def _is_complex_dtype(value_dtype) -> bool:
    return _is_complex32_dtype(value_dtype) or value_dtype in (torch.complex64, torch.complex128)


def _component_dtype_for_complex(value_dtype):
    if _is_complex32_dtype(value_dtype):
        return torch.float16
    if value_dtype == torch.complex64:
        return torch.float32
    if value_dtype == torch.complex128:
        return torch.float64
    raise TypeError(f"Unsupported complex dtype: {value_dtype}")


def _tolerance_for_dtype(value_dtype):
    """(atol, rtol) per value dtype, as _common.py:123-134."""
    if value_dtype == torch.float16:
        return 2e-3, 2e-3
    if _is_complex32_dtype(value_dtype):
        return 5e-3, 5e-3
    if value_dtype == torch.bfloat16:
        return 1e-1, 1e-1
    if value_dtype in (torch.float32, torch.complex64):
        return 1e-6, 1e-5
    if value_dtype in (torch.float64, torch.complex128):
        return 1e-10, 1e-8
    return 1e-6, 1e-5


def _build_random_dense(dense_size, value_dtype, device):
    """Standard-normal values (complex: independent re / im), _common.py:203-214."""
    if value_dtype in (torch.float16, torch.bfloat16, torch.float32, torch.float64):
        return torch.randn(dense_size, dtype=value_dtype, device=device)
    if _is_complex32_dtype(value_dtype):
        pair = torch.randn((dense_size, 2), dtype=torch.float16, device=device)
        return torch.view_as_complex(pair)
    if _is_complex_dtype(value_dtype):
        comp = _component_dtype_for_complex(value_dtype)
        return torch.complex(torch.randn(dense_size, dtype=comp, device=device),
                             torch.randn(dense_size, dtype=comp, device=device))
    raise TypeError(f"Unsupported value dtype: {value_dtype}")


def _build_random_csr(n_rows, n_cols, nnz, value_dtype, index_dtype, device):
    """Random CSR as in _common.py:223-238: uniformly drawn rows (sorted, so
    indptr is a bincount prefix), columns drawn with replacement and left
    unsorted, normal values; indptr is int64."""
    if nnz <= 0 or n_rows <= 0 or n_cols <= 0:
        return (torch.empty(0, dtype=value_dtype, device=device),
                torch.empty(0, dtype=index_dtype, device=device),
                torch.zeros(n_rows + 1, dtype=torch.int64, device=device))
    rows = torch.sort(torch.randint(0, n_rows, (nnz,), device=device)).values
    indptr = torch.zeros(n_rows + 1, dtype=torch.int64, device=device)
    indptr[1:] = torch.cumsum(torch.bincount(rows, minlength=n_rows), dim=0)
    indices = torch.randint(0, n_cols, (nnz,), dtype=index_dtype, device=device)
    return _build_random_dense(nnz, value_dtype, device), indices, indptr


def _benchmark_cuda_op(op, warmup, iters):
    """(last output, mean ms per call) with a device sync around the timed
    loop, as _common.py:353-371."""
    warmup = max(0, int(warmup))
    iters = max(1, int(iters))
    out = None
    for _ in range(warmup):
        out = op()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(iters):
        out = op()
    torch.cuda.synchronize()
    return out, (time.perf_counter() - t0) * 1000.0 / iters


__all__ = (
    "SUPPORTED_VALUE_DTYPES",
    "SUPPORTED_INDEX_DTYPES",
    "_INDEX_LIMIT_INT32",
    "_torch_complex32_dtype",
    "_is_complex32_dtype",
    "_is_complex_dtype",
    "_component_dtype_for_complex",
    "_tolerance_for_dtype",
    "_build_random_dense",
    "_build_random_csr",
    "_benchmark_cuda_op",
    "cp",
    "cpx_sparse",
    "time",
    "torch",
)
