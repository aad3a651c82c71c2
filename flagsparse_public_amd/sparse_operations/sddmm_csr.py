"""CSR SDDMM behind the reference's operator interface (sddmm_csr.py).

out[p] = alpha * dot(x[row(p)], y[col(p)]) + beta * data[p] over the entries
of a CSR pattern, fp32 / fp64.  Names, arguments, validation order and
messages, return tuples and ``meta`` keys follow sddmm_csr.py:9-420; the
kernel is the gfx950 ``fs_sddmm_csr`` (csrc/sddmm.hip: a lane group per entry,
16-byte row loads, DPP group reduction).  The reference's diagnostic variants
map to accumulation modes: "baseline" / "altreduce" accumulate in the value
dtype, "acc64" in fp64, "acc64_out64" in fp64 with an fp64 output.
``allow_fallback=True`` keeps the reference's behaviour -- a failed launch is
replaced by the torch computation on the same device and reported in meta --
but a missing library is never hidden (it raises before any launch).
"""

from __future__ import annotations

import time

import torch

from .. import _lib
from ._common import _benchmark_cuda_op, _build_random_csr, _build_random_dense, _tolerance_for_dtype

SUPPORTED_SDDMM_VALUE_DTYPES = (torch.float32, torch.float64)
SUPPORTED_SDDMM_DIAGNOSTIC_VARIANTS = ("baseline", "acc64", "acc64_out64", "altreduce")


class SDDMMPrepared:
    """Prepared CSR pattern metadata (sddmm_csr.py:9-33)."""

    __slots__ = ("indices", "indptr", "shape", "n_rows", "n_cols", "nnz", "row_ids", "block_k",
                 "num_warps")

    def __init__(self, indices, indptr, shape, row_ids, block_k, num_warps):
        self.indices = indices
        self.indptr = indptr
        self.shape = (int(shape[0]), int(shape[1]))
        self.n_rows = self.shape[0]
        self.n_cols = self.shape[1]
        self.nnz = int(indices.numel())
        self.row_ids = row_ids
        self.block_k = int(block_k)
        self.num_warps = int(num_warps)


def _resolve_sddmm_launch_config(k):
    """(block_k, num_warps) reported in meta (sddmm_csr.py:36-43)."""
    if k <= 32:
        return 32, 2
    if k <= 64:
        return 64, 4
    if k <= 128:
        return 64, 4
    return 128, 8


def _prepare_sddmm_csr_pattern(indices, indptr, shape):
    """Pattern checks of sddmm_csr.py:46-79, same order and messages."""
    if len(shape) != 2:
        raise ValueError("shape must be a 2-tuple")
    if indices.ndim != 1 or indptr.ndim != 1:
        raise ValueError("indices and indptr must be 1D tensors")
    n_rows, n_cols = int(shape[0]), int(shape[1])
    if n_rows < 0 or n_cols < 0:
        raise ValueError("shape dimensions must be non-negative")
    if indptr.numel() != n_rows + 1:
        raise ValueError(f"indptr length must be n_rows+1={n_rows + 1}, got {indptr.numel()}")
    if not indices.is_cuda or not indptr.is_cuda:
        raise ValueError("indices and indptr must be CUDA tensors")
    if indices.dtype != torch.int32:
        raise TypeError("indices dtype must be torch.int32")
    if indptr.dtype not in (torch.int32, torch.int64):
        raise TypeError("indptr dtype must be torch.int32 or torch.int64")
    indptr64 = indptr.to(torch.int64).contiguous()
    indices = indices.contiguous()
    nnz = int(indices.numel())
    if indptr64.numel() > 0 and int(indptr64[0].item()) != 0:
        raise ValueError("indptr[0] must be 0")
    if indptr64.numel() > 0 and int(indptr64[-1].item()) != nnz:
        raise ValueError(f"indptr[-1] must equal nnz={nnz}")
    if indptr64.numel() > 1 and bool(torch.any(indptr64[1:] < indptr64[:-1]).item()):
        raise ValueError("indptr must be nondecreasing")
    if nnz > 0:
        if int(indices.min().item()) < 0 or int(indices.max().item()) >= n_cols:
            raise IndexError("indices out of range for shape[1]")
    return indices, indptr64, (n_rows, n_cols)


def _build_row_ids(indptr):
    n_rows = int(indptr.numel()) - 1
    if n_rows <= 0:
        return torch.empty(0, dtype=torch.int32, device=indptr.device)
    return torch.repeat_interleave(torch.arange(n_rows, dtype=torch.int32, device=indptr.device),
                                   indptr[1:] - indptr[:-1])


def prepare_sddmm_csr(indices, indptr, shape, k_hint=64):
    """Validated pattern + per-entry row ids (sddmm_csr.py:93-105)."""
    indices, indptr, shape = _prepare_sddmm_csr_pattern(indices, indptr, shape)
    block_k, num_warps = _resolve_sddmm_launch_config(int(k_hint))
    return SDDMMPrepared(indices=indices, indptr=indptr, shape=shape,
                         row_ids=_build_row_ids(indptr), block_k=block_k, num_warps=num_warps)


def _validate_sddmm_dense_inputs(data, prepared, x, y):
    """x / y / data checks of sddmm_csr.py:205-226."""
    if x.ndim != 2 or y.ndim != 2:
        raise ValueError("x and y must be 2D dense tensors")
    if not x.is_cuda or not y.is_cuda:
        raise ValueError("x and y must be CUDA tensors")
    if x.device != y.device or x.device != prepared.indices.device:
        raise ValueError("x, y, and sparse pattern must be on the same CUDA device")
    if x.dtype not in SUPPORTED_SDDMM_VALUE_DTYPES:
        raise TypeError("x dtype must be torch.float32 or torch.float64")
    if y.dtype != x.dtype:
        raise TypeError("y dtype must match x dtype")
    if data is not None and data.dtype != x.dtype:
        raise TypeError("data dtype must match x/y dtype")
    if x.shape[0] != prepared.n_rows:
        raise ValueError(f"x.shape[0] must be n_rows={prepared.n_rows}, got {x.shape[0]}")
    if y.shape[0] != prepared.n_cols:
        raise ValueError(f"y.shape[0] must be n_cols={prepared.n_cols}, got {y.shape[0]}")
    if x.shape[1] != y.shape[1]:
        raise ValueError("x and y must have the same K dimension")
    if data is not None and data.numel() != prepared.nnz:
        raise ValueError("data length must equal nnz of sparse pattern")
    return int(x.shape[1])


def _prepare_validated_sddmm_out(prepared, x, out, out_dtype=None):
    nnz = prepared.nnz
    target_dtype = x.dtype if out_dtype is None else out_dtype
    if out is None:
        return torch.empty(nnz, dtype=target_dtype, device=x.device)
    if out.ndim != 1 or out.numel() != nnz:
        raise ValueError("out must be a 1D tensor with length nnz")
    if not out.is_cuda or out.device != x.device:
        raise ValueError("out must be a CUDA tensor on the same device as x")
    if out.dtype != target_dtype:
        raise TypeError("out dtype must match the requested output dtype")
    return out


def _normalize_sddmm_diagnostic_variant(variant):
    if variant is None:
        return "baseline"
    variant = str(variant).strip().lower()
    if variant not in SUPPORTED_SDDMM_DIAGNOSTIC_VARIANTS:
        supported = ", ".join(SUPPORTED_SDDMM_DIAGNOSTIC_VARIANTS)
        raise ValueError(f"Unsupported SDDMM diagnostic variant {variant!r}; "
                         f"expected one of: {supported}")
    return variant


def _sddmm_reference(indices, indptr, x, y, data, alpha, beta):
    """torch computation of sddmm_csr.py:423-437 (the fallback's path)."""
    n_rows = int(indptr.numel()) - 1
    row_ids = torch.repeat_interleave(torch.arange(n_rows, dtype=torch.int64, device=indices.device),
                                      indptr[1:] - indptr[:-1])
    if row_ids.numel() == 0:
        return torch.empty(0, dtype=x.dtype, device=x.device)
    vals = alpha * torch.sum(x[row_ids] * y[indices.to(torch.int64)], dim=1)
    if data is not None:
        vals = vals + beta * data
    return vals


def _run_sddmm_prepared(prepared, x, y, data, alpha, beta, out, allow_fallback=False,
                        variant="baseline", out_dtype=None):
    variant = _normalize_sddmm_diagnostic_variant(variant)
    if out_dtype is None:
        out_dtype = torch.float64 if variant == "acc64_out64" else x.dtype
    out = _prepare_validated_sddmm_out(prepared, x, out, out_dtype=out_dtype)
    if prepared.nnz == 0:
        return out, {"block_k": prepared.block_k, "num_warps": prepared.num_warps,
                     "fallback_used": False, "variant": variant,
                     "acc_dtype": "float64" if out_dtype == torch.float64 else "float32",
                     "out_dtype": str(out_dtype).replace("torch.", "")}
    wide = variant in ("acc64", "acc64_out64")
    acc_name = "float64" if (wide or x.dtype == torch.float64) else "float32"
    meta = {"block_k": prepared.block_k, "num_warps": prepared.num_warps, "fallback_used": False,
            "variant": variant, "acc_dtype": acc_name,
            "out_dtype": str(out.dtype).replace("torch.", "")}
    k_dim = int(x.shape[1])
    meta["block_k"], meta["num_warps"] = _resolve_sddmm_launch_config(k_dim)
    _lib.load()  # a missing library raises here, before any fallback
    # the kernel writes entry p at element p: a strided out gets a dense
    # staging buffer, copied back
    target = out if out.is_contiguous() else torch.empty(out.shape, dtype=out.dtype,
                                                          device=out.device)
    csr = _lib.FsCsr(value_type=_lib.value_code(x.dtype), index_type=_lib.FS_I32,
                     ptr_type=_lib.FS_I64, reserved=0, n_rows=prepared.n_rows,
                     n_cols=prepared.n_cols, nnz=prepared.nnz, row_ptr=_lib.ptr(prepared.indptr),
                     col_idx=_lib.ptr(prepared.indices),
                     values=_lib.ptr(data) if data is not None else None)
    try:
        _lib.sddmm(csr, prepared.row_ids, x, y, k_dim, target, _lib.value_code(out_dtype),
                   _lib.FS_ACC_WIDE if wide else _lib.FS_ACC_NATIVE, alpha,
                   beta if data is not None else 0.0, x.device)
        if target is not out:
            out.copy_(target)
    except Exception:
        if not allow_fallback:
            raise
        out.copy_(_sddmm_reference(prepared.indices, prepared.indptr, x, y, data, alpha,
                                   beta).to(out.dtype))
        meta["fallback_used"] = True
    return out, meta


def flagsparse_sddmm_csr(data=None, indices=None, indptr=None, x=None, y=None, shape=None,
                         alpha=1.0, beta=0.0, prepared=None, out=None, return_time=False,
                         return_meta=False, allow_fallback=False):
    """CSR SDDMM: out[p] = alpha * dot(x[row(p)], y[col(p)]) + beta * data[p]
    (sddmm_csr.py:333-420)."""
    prepare_ms = 0.0
    if prepared is None:
        if any(v is None for v in (indices, indptr, shape)):
            raise ValueError("indices, indptr, and shape are required when prepared is not provided")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        k_hint = int(x.shape[1]) if (x is not None and x.ndim == 2) else 64
        prepared = prepare_sddmm_csr(indices, indptr, shape, k_hint=k_hint)
        torch.cuda.synchronize()
        prepare_ms = (time.perf_counter() - t0) * 1000.0
    elif not isinstance(prepared, SDDMMPrepared):
        raise TypeError("prepared must be a SDDMMPrepared instance")
    if x is None or y is None:
        raise ValueError("x and y are required")
    if data is None and float(beta) != 0.0:
        raise ValueError("data is required when beta is non-zero")
    k_dim = _validate_sddmm_dense_inputs(data, prepared, x, y)
    if k_dim == 0:
        out = _prepare_validated_sddmm_out(prepared, x, out)
        if beta == 0.0 or data is None:
            out.zero_()
        else:
            out.copy_(data * beta)
        meta = {"prepare_ms": prepare_ms, "block_k": prepared.block_k,
                "num_warps": prepared.num_warps, "fallback_used": False}
        if return_time and return_meta:
            return out, 0.0, meta
        if return_time:
            return out, 0.0
        if return_meta:
            return out, meta
        return out
    t0 = None
    if return_time:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
    out_tensor, launch_meta = _run_sddmm_prepared(
        prepared, x.contiguous(), y.contiguous(),
        data.contiguous() if data is not None else None, alpha, beta, out,
        allow_fallback=allow_fallback)
    elapsed_ms = None
    if return_time:
        torch.cuda.synchronize()
        elapsed_ms = (time.perf_counter() - t0) * 1000.0
    meta = {"prepare_ms": prepare_ms, **launch_meta}
    if return_time and return_meta:
        return out_tensor, elapsed_ms, meta
    if return_time:
        return out_tensor, elapsed_ms
    if return_meta:
        return out_tensor, meta
    return out_tensor


def benchmark_sddmm_case(n_rows=1024, n_cols=1024, nnz=16384, k_dim=64, value_dtype=torch.float32,
                         warmup=10, iters=30, alpha=1.0, beta=0.0, run_cusparse=False):
    """SDDMM timed beside the torch sampled-dot computation and, with
    run_cusparse, rocSPARSE's SDDMM; result keys of sddmm_csr.py:472-571."""
    if value_dtype not in SUPPORTED_SDDMM_VALUE_DTYPES:
        raise TypeError("value_dtype must be torch.float32 or torch.float64")
    device = torch.device("cuda")
    data, indices, indptr = _build_random_csr(n_rows, n_cols, nnz, value_dtype, torch.int32, device)
    x = _build_random_dense((n_rows, k_dim), value_dtype, device)
    y = _build_random_dense((n_cols, k_dim), value_dtype, device)
    prepared = prepare_sddmm_csr(indices, indptr, (n_rows, n_cols), k_hint=k_dim)

    def op():
        return flagsparse_sddmm_csr(data=data, x=x, y=y, alpha=alpha, beta=beta,
                                    prepared=prepared, return_time=False)

    values, ms = _benchmark_cuda_op(op, warmup=warmup, iters=iters)
    ref_values, pytorch_ms = _benchmark_cuda_op(
        lambda: _sddmm_reference(indices, indptr.to(torch.int64), x, y, data, alpha, beta),
        warmup=warmup, iters=iters)
    atol, rtol = _tolerance_for_dtype(value_dtype)
    match = bool(torch.allclose(values, ref_values, atol=atol, rtol=rtol))
    max_abs = float(torch.max(torch.abs(values - ref_values)).item()) if values.numel() > 0 else 0.0
    vendor_ms = vendor_match = vendor_values = None
    reason = None
    if run_cusparse:
        # the vendor column: rocSPARSE's SDDMM in the slot of the reference's
        # CuPy sampled dot (sddmm_csr.py:515-538)
        from .. import vendor
        if not vendor.available():
            reason = "rocSPARSE baseline unavailable"
        elif prepared.nnz == 0:
            reason = "empty pattern"
        else:
            try:
                vout = torch.empty(prepared.nnz, dtype=value_dtype, device=device)
                bound = vendor.VendorSddmm(prepared.indices, prepared.indptr, prepared.shape,
                                           x.contiguous(), y.contiguous(), vout)

                def vrun():
                    if beta != 0.0:
                        vout.copy_(data)
                    bound.run(alpha, beta)
                    return vout
                vendor_values, vendor_ms = _benchmark_cuda_op(vrun, warmup=warmup, iters=iters)
                vendor_match = bool(torch.allclose(vendor_values, ref_values, atol=atol,
                                                   rtol=rtol))
            except Exception as exc:  # reported in backend_status like the reference
                reason = str(exc)
    return {
        "parameters": {"n_rows": n_rows, "n_cols": n_cols, "nnz": nnz, "k_dim": k_dim,
                       "value_dtype": str(value_dtype), "warmup": warmup, "iters": iters,
                       "alpha": alpha, "beta": beta},
        "performance": {"triton_ms": ms, "pytorch_ms": pytorch_ms, "cusparse_ms": vendor_ms,
                        "triton_speedup_vs_pytorch": (pytorch_ms / ms if ms > 0 else None),
                        "triton_speedup_vs_cusparse": (vendor_ms / ms if (vendor_ms and ms > 0)
                                                       else None)},
        "verification": {"triton_match_pytorch": match, "triton_max_abs_error": max_abs,
                         "cusparse_match_pytorch": vendor_match},
        "backend_status": {"cusparse_unavailable_reason": reason,
                           "vendor_library": "rocSPARSE"},
        "samples": {"triton": values, "pytorch": ref_values, "cusparse": vendor_values},
    }


__all__ = ["SUPPORTED_SDDMM_VALUE_DTYPES", "benchmark_sddmm_case", "SDDMMPrepared", "prepare_sddmm_csr",
           "flagsparse_sddmm_csr"]
