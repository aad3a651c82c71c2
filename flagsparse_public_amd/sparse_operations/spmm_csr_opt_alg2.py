"""Hardware-aware CSR SpMM "opt-alg2" entry points on gfx950.

Mirrors ``flagsparse/sparse_operations/spmm_csr_opt_alg2.py``: the alg2
bucket table and launch-resolution rules are kept (they feed ``return_meta``),
device properties are read from the live GPU (on MI355X: 256 CUs, wave64),
and the multiply itself is the HIP row-split SpMM with native accumulation.
``return_meta`` additionally reports the HIP schedule that ran.
"""

from __future__ import annotations

import math
import time

import torch

from .. import _lib
from ..rowsplit import Schedule
from ..schedule_choice import choose_schedule
from ._common import _INDEX_LIMIT_INT32
from .spmm_csr import (  # noqa: F401  (re-exported like the reference)
    _launch_spmm,
    _plan_for,
    _prepare_spmm_csr_matrix,
    _spmm_coo_reference_tolerance,
    _spmm_target,
    _spmm_validation_metrics,
    flagsparse_spmm_csr,
    flagsparse_spmm_csr_opt,
    prepare_spmm_csr_opt,
    spmm_launch_schedule,
)

SUPPORTED_SPMM_OPT_ALG2_DTYPES = (torch.float32, torch.float64)


class PreparedCsrSpmmOptAlg2:
    """Cached metadata for alg2 calls (spmm_csr_opt_alg2.py:20-57)."""

    __slots__ = ("data", "kernel_indices", "kernel_indptr", "shape", "n_rows", "n_cols",
                 "row_lengths", "max_row_nnz", "opt_buckets", "supports_opt", "schedule",
                 "library_schedule", "_plans", "_csr")

    def __init__(self, data, kernel_indices, kernel_indptr, shape, n_rows, n_cols, row_lengths,
                 max_row_nnz, opt_buckets, schedule=None, library_schedule=False):
        self.data = data
        self.kernel_indices = kernel_indices
        self.kernel_indptr = kernel_indptr
        self.shape = (int(shape[0]), int(shape[1]))
        self.n_rows = int(n_rows)
        self.n_cols = int(n_cols)
        self.row_lengths = row_lengths
        self.max_row_nnz = int(max_row_nnz)
        self.opt_buckets = opt_buckets
        self.supports_opt = data.dtype in SUPPORTED_SPMM_OPT_ALG2_DTYPES
        self.schedule = schedule
        self.library_schedule = bool(library_schedule)  # chosen by the library, not the caller
        self._plans = {}
        self._csr = None

    def csr_struct(self):
        if self._csr is None:
            self._csr = _lib.make_csr(self.data, self.kernel_indices, self.kernel_indptr,
                                      self.n_rows, self.n_cols)
        return self._csr


def _spec(label, kind, lo, hi, batch, block_k, cap, segments=1):
    return {"label": label, "kind": kind, "min_row_nnz": lo, "max_row_nnz": hi,
            "batch_rows": batch, "block_k": block_k, "block_n_cap": cap, "segments": segments}


# Bucket tables of spmm_csr_opt_alg2.py:60-166.
_SPMM_OPT_ALG2_BUCKET_SPECS_F32 = (
    _spec("short_16", "batched2d", 0, 16, 8, 16, 32),
    _spec("short_64", "batched2d", 17, 64, 4, 32, 64),
    _spec("row_256", "row2d", 65, 256, 1, 64, 64),
    _spec("seg_1024", "row2d_segmented", 257, 1024, 1, 64, 128, 4),
    _spec("seg_long", "row2d_segmented", 1025, None, 1, 128, 128, 8),
)
_SPMM_OPT_ALG2_BUCKET_SPECS_F64 = (
    _spec("short_16", "batched2d", 0, 16, 4, 16, 16),
    _spec("short_64", "batched2d", 17, 64, 2, 32, 32),
    _spec("row_256", "row2d", 65, 256, 1, 32, 32),
    _spec("seg_1024", "row2d_segmented", 257, 1024, 1, 32, 64, 4),
    _spec("seg_long", "row2d_segmented", 1025, None, 1, 64, 64, 8),
)


def _spmm_opt_alg2_bucket_specs(dtype):
    return _SPMM_OPT_ALG2_BUCKET_SPECS_F64 if dtype == torch.float64 else _SPMM_OPT_ALG2_BUCKET_SPECS_F32


def _round_down_power_of_two(value):
    return 1 if value <= 1 else 1 << (int(value).bit_length() - 1)


def _normalize_spmm_opt_alg2_device_props(device):
    """Live device properties (spmm_csr_opt_alg2.py:175-195)."""
    props = torch.cuda.get_device_properties(device)
    return {
        "device_name": str(getattr(props, "name", "hip")),
        "warp_size": int(getattr(props, "warp_size", 64) or 64),
        "sm_count": int(getattr(props, "multi_processor_count", 0) or 0),
        "max_threads_per_mp": int(getattr(props, "max_threads_per_multi_processor", 2048) or 2048),
        "max_threads_per_block": int(getattr(props, "max_threads_per_block", 1024) or 1024),
        "shared_memory_per_block": int(getattr(props, "shared_memory_per_block", 0) or 0),
        "capability": (int(getattr(props, "major", 0) or 0), int(getattr(props, "minor", 0) or 0)),
    }


def _select_spmm_opt_alg2_block_n(n_dense_cols, block_n_cap):
    """Column tile capped per bucket (spmm_csr_opt_alg2.py:198-209)."""
    block_n = 128
    for cap in (8, 16, 32, 64):
        if n_dense_cols <= cap:
            block_n = cap
            break
    return max(8, min(int(block_n), int(block_n_cap)))


def _select_spmm_opt_alg2_num_warps(bucket, block_n, device_props, dtype):
    """Wave count rule of spmm_csr_opt_alg2.py:212-249 (recorded in meta)."""
    warp = max(1, int(device_props["warp_size"]))
    max_tpb = max(32, int(device_props["max_threads_per_block"]))
    max_tpm = max(max_tpb, int(device_props["max_threads_per_mp"]))
    lane_need = max(1, math.ceil(int(block_n) / warp))
    kind, block_k = bucket["kind"], int(bucket["block_k"])
    segments, batch = int(bucket.get("segments", 1)), int(bucket.get("batch_rows", 1))
    if kind == "batched2d":
        desired = max(lane_need, 1 if block_k <= 16 else 2)
        if batch >= 8 and block_n >= 64:
            desired = max(desired, 2)
    elif kind == "row2d":
        desired = max(lane_need, 2 if block_k <= 32 else 4)
    else:
        desired = max(lane_need, 4)
        if segments >= 4:
            desired = max(desired, 8 if block_k >= 64 else 4)
        if segments >= 8:
            desired = max(desired, 16 if dtype == torch.float32 and block_n >= 64 else 8)
    if dtype == torch.float64 and desired > 8:
        desired = 8
    if max_tpm < 1536 and desired > 8:
        desired = 8
    max_supported = min(16, max(1, max_tpb // warp), max(1, max_tpm // warp))
    supported = [v for v in (1, 2, 4, 8, 16) if v <= max_supported]
    if not supported:
        return 1
    clipped = min(desired, supported[-1])
    for v in reversed(supported):
        if v <= clipped:
            return v
    return supported[0]


def _select_spmm_opt_alg2_num_stages(bucket, block_n, num_warps, device_props, dtype):
    """Stage rule of spmm_csr_opt_alg2.py:252-275 (recorded in meta)."""
    kind, segments = bucket["kind"], int(bucket.get("segments", 1))
    if dtype == torch.float64:
        stages = 1 if kind == "row2d_segmented" else (2 if block_n <= 32 and num_warps <= 4 else 1)
    elif kind == "batched2d":
        stages = 2
    elif kind == "row2d":
        stages = 2 if block_n <= 64 else 1
    else:
        stages = 1 if segments >= 8 or num_warps >= 16 else 2
    smem = int(device_props["shared_memory_per_block"])
    if smem and smem < 65536:
        stages = min(stages, 2)
    if kind == "row2d_segmented" and block_n >= 128:
        stages = min(stages, 2)
    return max(1, min(int(stages), 4))


def _resolve_spmm_opt_alg2_launch(bucket, n_dense_cols, dtype, device_props):
    block_n = _select_spmm_opt_alg2_block_n(n_dense_cols, bucket["block_n_cap"])
    num_warps = _select_spmm_opt_alg2_num_warps(bucket, block_n, device_props, dtype)
    num_stages = _select_spmm_opt_alg2_num_stages(bucket, block_n, num_warps, device_props, dtype)
    return {"bucket_label": bucket["label"], "kind": bucket["kind"],
            "block_k": int(bucket["block_k"]), "block_n": int(block_n),
            "num_warps": int(num_warps), "num_stages": int(num_stages),
            "batch_rows": int(bucket.get("batch_rows", 1)),
            "segments": int(bucket.get("segments", 1)), "row_count": int(bucket["rows"].numel())}


def _build_spmm_opt_alg2_buckets(row_lengths, dtype):
    """Row lists per alg2 bucket, inclusive bounds (spmm_csr_opt_alg2.py:307-335)."""
    device = row_lengths.device
    count = int(row_lengths.numel())
    idx_dtype = torch.int32 if count <= _INDEX_LIMIT_INT32 else torch.int64
    all_rows = torch.arange(count, device=device, dtype=idx_dtype)
    buckets = []
    for spec in _spmm_opt_alg2_bucket_specs(dtype):
        lo, hi = int(spec["min_row_nnz"]), spec["max_row_nnz"]
        if hi is None:
            mask = row_lengths >= lo
        elif lo <= 0:
            mask = row_lengths <= hi
        else:
            mask = (row_lengths >= lo) & (row_lengths <= hi)
        rows = all_rows[mask]
        if rows.numel() == 0:
            continue
        b = dict(spec)
        b["rows"] = rows
        buckets.append(b)
    return buckets


def prepare_spmm_csr_opt_alg2(data, indices, indptr, shape, schedule=None):
    """Validate once and cache metadata (spmm_csr_opt_alg2.py:632-653)."""
    stats_out = []
    (data, kernel_indices, kernel_indptr, n_rows, n_cols, row_lengths,
     max_row_nnz) = _prepare_spmm_csr_matrix(data, indices, indptr, shape, _stats_out=stats_out)
    library = not isinstance(schedule, Schedule)
    schedule, _ = choose_schedule("spmm64", schedule, data, n_rows, n_cols, stats_out[0])
    return PreparedCsrSpmmOptAlg2(
        data=data, kernel_indices=kernel_indices, kernel_indptr=kernel_indptr, shape=shape,
        n_rows=n_rows, n_cols=n_cols, row_lengths=row_lengths, max_row_nnz=max_row_nnz,
        opt_buckets=_build_spmm_opt_alg2_buckets(row_lengths, data.dtype), schedule=schedule,
        library_schedule=library)


def _validate_spmm_opt_alg2_runtime_inputs(prepared, B, out):
    """Runtime checks of spmm_csr_opt_alg2.py:656-677."""
    if not prepared.supports_opt:
        raise TypeError("flagsparse_spmm_csr_opt_alg2 only supports float32 and float64")
    if B is None:
        raise ValueError("B is required")
    if B.ndim != 2:
        raise ValueError("B must be a 2D dense tensor")
    if not B.is_cuda:
        raise ValueError("B must be a CUDA tensor")
    if B.device != prepared.data.device:
        raise ValueError("B must be on the same CUDA device as sparse matrix data")
    if B.dtype != prepared.data.dtype:
        raise TypeError("B dtype must match sparse matrix dtype")
    if B.shape[0] != prepared.n_cols:
        raise ValueError(f"B.shape[0] must be n_cols={prepared.n_cols}, got {B.shape[0]}")
    if out is not None:
        if not out.is_cuda:
            raise ValueError("out must be a CUDA tensor")
        if out.device != prepared.data.device:
            raise ValueError("out must be on the same CUDA device as sparse matrix data")
        if tuple(out.shape) != (prepared.n_rows, int(B.shape[1])) or out.dtype != prepared.data.dtype:
            raise ValueError("out shape/dtype must match result")


def _hip_spmm_csr_impl_opt_alg2_prepared(prepared, B, return_meta=False, out=None, alpha=1.0,
                                         beta=0.0, schedule=None):
    """Run the native-accumulation SpMM; meta as spmm_csr_opt_alg2.py:771-798."""
    device_props = _normalize_spmm_opt_alg2_device_props(prepared.data.device)
    n_dense = int(B.shape[1])
    sch, col_tile = spmm_launch_schedule(prepared, n_dense, schedule)
    plan = _plan_for(prepared._plans, prepared.kernel_indptr, prepared.n_rows,
                     prepared.data.dtype, _lib.FS_ACC_NATIVE, sch, prepared.row_lengths)
    C = _launch_spmm(prepared.csr_struct(), plan, B, prepared.n_rows, n_dense, B.dtype, B.device,
                     out, alpha, beta, col_tile)
    meta = None
    if return_meta:
        launches = [_resolve_spmm_opt_alg2_launch(b, n_dense, prepared.data.dtype, device_props)
                    for b in prepared.opt_buckets]
        meta = {
            "device_name": device_props["device_name"],
            "sm_count": device_props["sm_count"],
            "warp_size": device_props["warp_size"],
            "bucket_hits": [{"bucket_label": l["bucket_label"], "kind": l["kind"],
                             "row_count": l["row_count"]} for l in launches],
            "launch_configs": launches,
            "max_row_nnz": prepared.max_row_nnz,
            "n_dense_cols": n_dense,
            "hip_schedule": {"short_max": sch.short_max, "chunk_nnz": sch.chunk_nnz,
                             "n_long_rows": plan.n_long, "n_chunks": plan.n_chunks,
                             "col_tile": col_tile, "unroll": sch.unroll},
        }
    return C, meta


_triton_spmm_csr_impl_opt_alg2_prepared = _hip_spmm_csr_impl_opt_alg2_prepared


def flagsparse_spmm_csr_opt_alg2(data=None, indices=None, indptr=None, B=None, shape=None,
                                 prepared=None, out=None, return_time=False, return_meta=False,
                                 alpha=1.0, beta=0.0):
    """Alg2 CSR @ dense (spmm_csr_opt_alg2.py:801-850); returns C, (C, ms),
    (C, meta) or (C, ms, meta)."""
    if prepared is not None and not isinstance(prepared, PreparedCsrSpmmOptAlg2):
        raise TypeError("prepared must be a PreparedCsrSpmmOptAlg2 instance")
    if prepared is None:
        if any(a is None for a in (data, indices, indptr, shape)):
            raise ValueError(
                "data, indices, indptr, and shape are required when prepared is not provided")
        prepared = prepare_spmm_csr_opt_alg2(data, indices, indptr, shape)
    elif shape is not None:
        sh = (int(shape[0]), int(shape[1]))
        if sh != prepared.shape:
            raise ValueError(f"shape {sh} does not match prepared.shape {prepared.shape}")
    _validate_spmm_opt_alg2_runtime_inputs(prepared, B, out)
    if float(beta) != 0.0 and out is None:
        raise ValueError("beta != 0 requires out (the C that is scaled and accumulated)")
    B = B.contiguous()
    target = _spmm_target(out, beta)
    t0 = None
    if return_time:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
    C, meta = _triton_spmm_csr_impl_opt_alg2_prepared(prepared, B, return_meta=return_meta,
                                                      out=target, alpha=alpha, beta=beta)
    elapsed_ms = None
    if return_time:
        torch.cuda.synchronize()
        elapsed_ms = (time.perf_counter() - t0) * 1000.0
    if out is not None and C is not out:
        out.copy_(C)
        C = out
    if return_time and return_meta:
        return C, elapsed_ms, meta
    if return_time:
        return C, elapsed_ms
    if return_meta:
        return C, meta
    return C


def _spmm_opt_alg2_reference_error(candidate, reference, value_dtype):
    atol, rtol = _spmm_coo_reference_tolerance(value_dtype)
    if candidate.numel() == 0:
        return 0.0
    return float(torch.max(torch.abs(candidate - reference) /
                           (atol + rtol * torch.abs(reference))).item())


def benchmark_spmm_opt_alg2_case(n_rows=4096, n_cols=4096, nnz=65536, n_dense_cols=32,
                                 value_dtype=torch.float32, index_dtype=torch.int32, warmup=20,
                                 iters=200, run_cusparse=True):
    """Base, opt and alg2 SpMM against torch sparse and the vendor library
    (rocSPARSE in the ``cusparse_*`` fields); spmm_csr_opt_alg2.py:862-987."""
    from .. import vendor
    from ._common import _benchmark_cuda_op, _build_random_csr, _build_random_dense
    from .spmm_csr import flagsparse_spmm_csr, flagsparse_spmm_csr_opt, prepare_spmm_csr_opt
    if value_dtype not in SUPPORTED_SPMM_OPT_ALG2_DTYPES:
        raise TypeError("benchmark_spmm_opt_alg2_case only supports float32 and float64")
    device = torch.device("cuda")
    data, indices, indptr = _build_random_csr(n_rows, n_cols, nnz, value_dtype, index_dtype, device)
    B = _build_random_dense((n_cols, n_dense_cols), value_dtype, device)
    shape = (n_rows, n_cols)
    prepared_opt = prepare_spmm_csr_opt(data, indices, indptr, shape)
    prepared_alg2 = prepare_spmm_csr_opt_alg2(data, indices, indptr, shape)
    base_values, base_ms = _benchmark_cuda_op(
        lambda: flagsparse_spmm_csr(data, indices, indptr, B, shape), warmup, iters)
    opt_values, opt_ms = _benchmark_cuda_op(
        lambda: flagsparse_spmm_csr_opt(B=B, prepared=prepared_opt), warmup, iters)
    alg2_values, alg2_ms = _benchmark_cuda_op(
        lambda: flagsparse_spmm_csr_opt_alg2(B=B, prepared=prepared_alg2), warmup, iters)
    ip64, ix64 = indptr.to(torch.int64), indices.to(torch.int64)
    ref_dtype = torch.float64 if value_dtype == torch.float32 else value_dtype
    csr_ref = torch.sparse_csr_tensor(ip64, ix64, data.to(ref_dtype), size=shape, device=device)
    torch_ref = torch.sparse.mm(csr_ref, B.to(ref_dtype)).to(value_dtype)
    torch_ms = None
    try:
        pt = torch.sparse_csr_tensor(ip64, ix64, data, size=shape, device=device)
        _, torch_ms = _benchmark_cuda_op(lambda: torch.sparse.mm(pt, B), warmup, iters)
    except Exception:
        torch_ms = None
    vendor_values = vendor_ms = vendor_reason = None
    if run_cusparse:
        _, run, vendor_reason = vendor.spmm_baseline(data, indices, indptr, shape, B)
        if run is not None:
            try:
                vendor_values, vendor_ms = _benchmark_cuda_op(run, warmup, iters)
                vendor_values = vendor_values.clone()
            except Exception as exc:
                vendor_reason = str(exc)

    def err(a, b):
        return _spmm_opt_alg2_reference_error(a, b, value_dtype) if b is not None else None

    return {
        "parameters": {"n_rows": n_rows, "n_cols": n_cols, "nnz": nnz,
                       "n_dense_cols": n_dense_cols, "value_dtype": str(value_dtype),
                       "index_dtype": str(index_dtype)},
        "performance": {"base_ms": base_ms, "opt_ms": opt_ms, "opt_alg2_ms": alg2_ms,
                        "torch_ms": torch_ms, "cusparse_ms": vendor_ms},
        "verification": {
            "base_vs_torch_err": err(base_values, torch_ref),
            "opt_vs_torch_err": err(opt_values, torch_ref),
            "opt_alg2_vs_torch_err": err(alg2_values, torch_ref),
            "base_vs_cusparse_err": err(base_values, vendor_values),
            "opt_vs_cusparse_err": err(opt_values, vendor_values),
            "opt_alg2_vs_cusparse_err": err(alg2_values, vendor_values),
        },
        "backend_status": {"cusparse_unavailable_reason": vendor_reason,
                           "vendor_library": "rocSPARSE"},
        "samples": {"base_triton": base_values, "opt_triton": opt_values,
                    "opt_alg2_triton": alg2_values, "torch_ref": torch_ref,
                    "cusparse_ref": vendor_values},
    }


__all__ = [
    "SUPPORTED_SPMM_OPT_ALG2_DTYPES",
    "benchmark_spmm_opt_alg2_case",
    "PreparedCsrSpmmOptAlg2",
    "prepare_spmm_csr_opt_alg2",
    "flagsparse_spmm_csr_opt_alg2",
]
