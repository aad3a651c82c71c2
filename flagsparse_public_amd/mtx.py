"""MatrixMarket (.mtx) input and the SuiteSparse batch harness of the SpMV.

``load_mtx_to_csr_torch`` mirrors the reference harness's loader
(tests/test_spmv.py:103-176): same reading rules, same return value
``(data, indices int64, indptr int64, (n_rows, n_cols))``; the parsing is the
native ``fs_mtx_read`` (csrc/mtx_io.cpp) instead of a Python dict per row.

``run_one_mtx`` / ``run_mtx_batch`` follow tests/test_spmv.py:416-665: load,
run the SpMV (prepared, timed with ``_benchmark_cuda_op``), compare with the
torch sparse reference in the promoted dtype and with the vendor library
(rocSPARSE, in the ``cusparse_ms`` / ``err_cu`` fields), and report one
result dict per file with the reference's keys and status values.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib

FIELDS = {0: "real", 1: "integer", 2: "complex", 3: "pattern"}
SYMMETRIES = {0: "general", 1: "symmetric", 2: "skew-symmetric", 3: "hermitian"}


def read_mtx(file_path):
    """(indptr, indices, values fp64) as numpy arrays, shape, info dict."""
    lib = _lib.load()
    info = _lib.FsMtxInfo()
    handle = ctypes.c_void_p()
    rc = lib.fs_mtx_read(str(file_path).encode(), ctypes.byref(info), ctypes.byref(handle))
    if rc != 0:
        raise ValueError(lib.fs_mtx_last_error().decode("utf-8", "replace"))
    try:
        indptr = np.empty(info.n_rows + 1, np.int64)
        indices = np.empty(info.nnz, np.int64)
        values = np.empty(info.nnz, np.float64)
        lib.fs_mtx_fill(handle, indptr.ctypes.data, indices.ctypes.data if info.nnz else None,
                        values.ctypes.data if info.nnz else None)
    finally:
        lib.fs_mtx_free(handle)
    meta = {"field": FIELDS[info.field], "symmetry": SYMMETRIES[info.symmetry]}
    return indptr, indices, values, (int(info.n_rows), int(info.n_cols)), meta


def load_mtx_to_csr_torch(file_path, dtype=torch.float32, device=None):
    """Load a MatrixMarket coordinate file as CSR torch tensors on `device`
    (tests/test_spmv.py:103-176): values summed in fp64 and cast to `dtype`,
    int64 indices and row pointer."""
    if device is None:
        device = torch.device("cuda" if torch.cuda.is_available() else "cpu")
    indptr, indices, values, shape, _ = read_mtx(file_path)
    data = torch.from_numpy(values).to(dtype=dtype, device=device)
    return (data, torch.from_numpy(indices).to(device), torch.from_numpy(indptr).to(device),
            shape)


def _tolerance(value_dtype):
    """(atol, rtol) of the reference's SuiteSparse harness (tests/test_spmv.py:406-413)."""
    if value_dtype in (torch.float16, torch.bfloat16):
        return 2e-3, 2e-3
    if value_dtype in (torch.float32, torch.complex64):
        return 1.25e-4, 1.25e-2
    if value_dtype in (torch.float64, torch.complex128):
        return 1e-12, 1e-10
    return 5e-3, 5e-3


def _random_vector(size, dtype, device):
    if dtype.is_complex:
        comp = torch.float32 if dtype == torch.complex64 else torch.float64
        return torch.complex(torch.randn(size, dtype=comp, device=device),
                             torch.randn(size, dtype=comp, device=device))
    return torch.randn(size, dtype=torch.float32, device=device).to(dtype)


def _err_ratio(actual, reference, atol, rtol):
    """max |d| / (atol + rtol |ref|) in fp64 (tests/test_spmv.py:179-184)."""
    if actual.numel() == 0:
        return 0.0
    diff = (actual.to(reference.dtype) - reference).abs().to(torch.float64)
    tol = (atol + rtol * reference.abs()).to(torch.float64)
    return float((diff / tol).max().item())


def run_one_mtx(mtx_path, value_dtype=torch.float32, index_dtype=torch.int32, warmup=10,
                iters=50, run_cusparse=True, block_nnz=256, max_segments=None, op="non"):
    """SpMV on one .mtx file: timings and errors (tests/test_spmv.py:416-615)."""
    from . import vendor
    from .sparse_operations._common import _benchmark_cuda_op
    from .sparse_operations.benchmarks import _torch_spmv_reference, _vendor_op_matrix
    from .sparse_operations.spmv_csr import (_normalize_spmv_op, _spmv_op_to_name,
                                             _spmv_op_transposes, flagsparse_spmv_csr,
                                             prepare_spmv_csr)
    op_code = _normalize_spmv_op(op)
    op_name = _spmv_op_to_name(op_code)
    device = torch.device("cuda")
    data, indices, indptr, shape = load_mtx_to_csr_torch(mtx_path, dtype=value_dtype,
                                                         device=device)
    indices = indices.to(index_dtype)
    n_rows, n_cols = shape
    transpose = _spmv_op_transposes(op_code)
    x = _random_vector(n_rows if transpose else n_cols, value_dtype, device)
    atol, rtol = _tolerance(value_dtype)
    prepared = prepare_spmv_csr(data, indices, indptr, shape, block_nnz=block_nnz,
                                max_segments=max_segments, op=op_code)
    y, hip_ms = _benchmark_cuda_op(lambda: flagsparse_spmv_csr(x=x, prepared=prepared),
                                   warmup, iters)
    ref_dtype = {torch.float16: torch.float32, torch.bfloat16: torch.float32,
                 torch.float32: torch.float64, torch.complex64: torch.complex128}.get(
                     value_dtype, value_dtype)
    ref = _torch_spmv_reference(data.to(ref_dtype), indices, indptr, x.to(ref_dtype), shape,
                                op_code)
    err_pt = _err_ratio(y, ref, atol, rtol)
    ok_pt = bool(torch.isfinite(torch.as_tensor(err_pt))) and err_pt <= 1.0
    vendor_ms = err_cu = None
    ok_cu = False
    if run_cusparse:
        vd, vi, vp, vshape = _vendor_op_matrix(data, indices, indptr, shape, op_code)
        _, run, _reason = vendor.spmv_baseline(vd, vi, vp, vshape, x, lambda v: v.clone())
        if run is not None:
            vy, vendor_ms = _benchmark_cuda_op(run, warmup, iters)
            err_cu = _err_ratio(y, vy.to(ref_dtype), atol, rtol)
            ok_cu = err_cu <= 1.0
    finite = bool(torch.isfinite(y.abs() if y.is_complex() else y).all().item())
    status = "PASS" if finite and (ok_pt or ok_cu) else "FAIL"
    return {
        "path": mtx_path, "shape": shape, "nnz": int(data.numel()), "error": None,
        "error_reason": None if status == "PASS" else (
            "non-finite FlagSparse output" if not finite else "tolerance exceeded"),
        "op": op_name, "transpose": transpose, "triton_ms": hip_ms, "cusparse_ms": vendor_ms,
        "pytorch_ms": None, "csc_ms": None, "err_pt": err_pt, "err_cu": err_cu,
        "triton_ok_pt": ok_pt, "triton_ok_cu": ok_cu, "status": status,
    }


def run_mtx_batch(mtx_paths, value_dtype=torch.float32, index_dtype=torch.int32, warmup=10,
                  iters=50, run_cusparse=True, on_result=None, op="non"):
    """run_one_mtx over many files; a file that fails gives status "ERROR"
    (tests/test_spmv.py:618-665)."""
    from .sparse_operations.spmv_csr import (_normalize_spmv_op, _spmv_op_to_name,
                                             _spmv_op_transposes)
    op_code = _normalize_spmv_op(op)
    results = []
    for path in mtx_paths:
        try:
            r = run_one_mtx(path, value_dtype=value_dtype, index_dtype=index_dtype,
                            warmup=warmup, iters=iters, run_cusparse=run_cusparse, op=op_code)
        except Exception as exc:
            r = {"path": path, "shape": (0, 0), "nnz": 0, "error": str(exc),
                 "error_reason": str(exc), "op": _spmv_op_to_name(op_code),
                 "transpose": _spmv_op_transposes(op_code), "triton_ms": None,
                 "cusparse_ms": None, "pytorch_ms": None, "csc_ms": None, "err_pt": None,
                 "err_cu": None, "triton_ok_pt": False, "triton_ok_cu": False,
                 "status": "ERROR"}
        results.append(r)
        if on_result is not None:
            on_result(r)
    return results


__all__ = ["load_mtx_to_csr_torch", "read_mtx", "run_one_mtx", "run_mtx_batch"]
