"""CSR SpMM on gfx950 behind the reference's operator interface.

Mirrors ``flagsparse/sparse_operations/spmm_csr.py``: ``flagsparse_spmm_csr``
(ALG1 default path), ``prepare_spmm_csr_opt`` / ``flagsparse_spmm_csr_opt``,
the error / tolerance helpers and the validation contract.  Kernels are the
HIP ones of ``csrc/spmm.hip`` (C ABI ``fs_spmm_csr``).

Numerics: the ALG1 default path accumulates fp32 in fp64 and rounds once
(spmm_csr.py:544-551 casts data and B to fp64), fp16 / bf16 in fp32; the opt
path accumulates fp32 / fp64 natively (chunked long rows of fp32 with
compensated summation).  Extensions: trailing ``alpha`` / ``beta``.
"""

from __future__ import annotations

import dataclasses
import time

import torch

from .. import _lib
from ..rowsplit import RowSplitPlan, Schedule, default_spmm_schedule
from ..schedule_choice import choose_schedule
from ._common import (
    SUPPORTED_INDEX_DTYPES,
    SUPPORTED_VALUE_DTYPES,
    _INDEX_LIMIT_INT32,
    _build_random_csr,
    _build_random_dense,
    _benchmark_cuda_op,
    _is_complex_dtype,
    _tolerance_for_dtype,
)

SUPPORTED_SPMM_VALUE_DTYPES = SUPPORTED_VALUE_DTYPES


# ------------------------------------------------------------ error metrics

def _spmm_relative_threshold(value_dtype):
    """Max-relative-error threshold per dtype (spmm_csr.py:6-15)."""
    if value_dtype == torch.float16:
        return 5e-3
    if value_dtype == torch.bfloat16:
        return 1e-2
    if value_dtype in (torch.float32, torch.complex64):
        return 1e-6
    if value_dtype in (torch.float64, torch.complex128):
        return 1e-12
    return 1e-6


def _spmm_coo_reference_tolerance(value_dtype):
    """(atol, rtol) of the opt reference check (spmm_csr.py:18-27)."""
    if value_dtype == torch.float16:
        return 2e-3, 2e-3
    if value_dtype == torch.bfloat16:
        return 1e-1, 1e-1
    if value_dtype in (torch.float32, torch.complex64):
        return 1e-4, 1e-2
    if value_dtype in (torch.float64, torch.complex128):
        return 1e-12, 1e-10
    return 1e-6, 1e-5


def _spmm_error_metrics(candidate, reference):
    """max_abs, max_rel (= max_abs / max|ref|), sum_rel (spmm_csr.py:30-74);
    complex values are compared by magnitude as in the reference."""
    if candidate.shape != reference.shape:
        raise ValueError("candidate and reference must have the same shape, got "
                         f"{candidate.shape} vs {reference.shape}")
    if candidate.numel() == 0:
        return {"max_abs_error": 0.0, "max_relative_error": 0.0, "sum_relative_error": 0.0,
                "reference_max_magnitude": 0.0, "reference_sum_magnitude": 0.0}
    if _is_complex_dtype(reference.dtype):
        ref_mag = torch.abs(reference)
        diff = torch.abs(torch.abs(candidate) - ref_mag)
    else:
        ref_mag = torch.abs(reference)
        diff = torch.abs(candidate - reference)
    max_abs = float(torch.max(diff).item())
    ref_max = float(torch.max(ref_mag).item())
    sum_abs = float(torch.sum(diff).item())
    ref_sum = float(torch.sum(ref_mag).item())

    def ratio(num, den):
        if den == 0.0:
            return 0.0 if num == 0.0 else float("inf")
        return num / den

    return {"max_abs_error": max_abs, "max_relative_error": ratio(max_abs, ref_max),
            "sum_relative_error": ratio(sum_abs, ref_sum), "reference_max_magnitude": ref_max,
            "reference_sum_magnitude": ref_sum}


def _spmm_validation_metrics(candidate, reference):
    """Metrics plus threshold / allclose verdicts (spmm_csr.py:77-90)."""
    m = _spmm_error_metrics(candidate, reference)
    thr = _spmm_relative_threshold(reference.dtype)
    atol, rtol = _tolerance_for_dtype(reference.dtype)
    m.update({"relative_threshold": thr, "matches_threshold": m["max_relative_error"] <= thr,
              "strict_allclose_match": torch.allclose(candidate, reference, atol=atol, rtol=rtol)})
    return m


# --------------------------------------------------------------- validation

def _prepare_spmm_csr_matrix(data, indices, indptr, shape, _stats_out=None):
    """Checks and messages of spmm_csr.py:197-255, device checks fused in one
    fs_csr_analyze pass.  Column indices become int32 and indptr int64 as in
    the reference (:247-248)."""
    if len(shape) != 2:
        raise ValueError("shape must be a 2-tuple: (n_rows, n_cols)")
    if data.ndim != 1 or indices.ndim != 1 or indptr.ndim != 1:
        raise ValueError("data, indices, and indptr must be 1D tensors")
    n_rows, n_cols = int(shape[0]), int(shape[1])
    if n_rows < 0 or n_cols < 0:
        raise ValueError("shape dimensions must be non-negative")
    if indptr.numel() != n_rows + 1:
        raise ValueError(f"indptr length must be n_rows+1={n_rows + 1}, got {indptr.numel()}")
    if data.numel() != indices.numel():
        raise ValueError("data and indices must have the same length (nnz)")
    if not all(t.is_cuda for t in (data, indices, indptr)):
        raise ValueError("data, indices, and indptr must be CUDA tensors")
    if not all(t.device == data.device for t in (indices, indptr)):
        raise ValueError("data, indices, and indptr must be on the same CUDA device")
    if data.dtype not in SUPPORTED_SPMM_VALUE_DTYPES:
        raise TypeError("data dtype must be one of: float16, bfloat16, float32, float64, "
                        "complex64, complex128")
    if indices.dtype not in SUPPORTED_INDEX_DTYPES:
        raise TypeError("indices dtype must be torch.int32 or torch.int64")
    if indptr.dtype not in SUPPORTED_INDEX_DTYPES:
        raise TypeError("indptr dtype must be torch.int32 or torch.int64")
    nnz = data.numel()
    data = data.contiguous()
    indices = indices.contiguous()
    indptr = indptr.contiguous()
    stats_dev = _lib.analyze_device(_lib.make_csr(data, indices, indptr, n_rows, n_cols), data.device)
    stats = stats_dev.cpu().tolist()
    if _stats_out is not None:
        _stats_out.append(stats_dev)
    if stats[0] != 0:
        raise ValueError("indptr[0] must be 0")
    if stats[1] != nnz:
        raise ValueError(f"indptr[-1] must equal nnz={nnz}, got {int(stats[1])}")
    if stats[2] > 0:
        raise ValueError("indptr must be nondecreasing")
    if nnz > 0:
        if stats[3] < 0 or stats[4] >= n_cols:
            raise IndexError("indices out of range for n_cols")
        if stats[4] > _INDEX_LIMIT_INT32:
            raise ValueError("column indices exceed the int32 range supported by the Triton kernel")
    kernel_indices = indices.to(torch.int32) if indices.dtype == torch.int64 else indices
    kernel_indptr = indptr.to(torch.int64)
    row_lengths = (kernel_indptr[1:] - kernel_indptr[:-1]) if n_rows > 0 \
        else kernel_indptr.new_empty((0,))
    max_row_nnz = int(stats[5]) if n_rows > 0 else 0
    return data, kernel_indices, kernel_indptr, n_rows, n_cols, row_lengths, max_row_nnz


def _prepare_spmm_csr_inputs(data, indices, indptr, B, shape):
    """Matrix checks plus B checks (spmm_csr.py:258-281)."""
    if B.ndim != 2:
        raise ValueError("B must be a 2D dense tensor")
    (data, kernel_indices, kernel_indptr, n_rows, n_cols, row_lengths,
     max_row_nnz) = _prepare_spmm_csr_matrix(data, indices, indptr, shape)
    if B.shape[0] != n_cols:
        raise ValueError(f"B.shape[0] must be n_cols={n_cols}, got {B.shape[0]}")
    if not B.is_cuda:
        raise ValueError("B must be a CUDA tensor")
    if B.device != data.device:
        raise ValueError("B must be on the same CUDA device as the sparse matrix")
    if B.dtype != data.dtype:
        raise TypeError("B dtype must match data dtype")
    return (data, kernel_indices, kernel_indptr, B.contiguous(), n_rows, n_cols,
            int(B.shape[1]), row_lengths, max_row_nnz)


def _select_spmm_alg1_warp_and_factor(n_dense_cols):
    """ALG1 (warp, factor) table of spmm_csr.py:284-296 (recorded, and used to
    resolve block_n defaults)."""
    if n_dense_cols > 64:
        return 32, 4
    if n_dense_cols > 32:
        return 32, 2
    if n_dense_cols > 16:
        return 32, 1
    if n_dense_cols > 8:
        return 16, 1
    if n_dense_cols > 4:
        return 8, 1
    return 4, 1


def _resolve_spmm_alg1_launch_config(n_dense_cols, max_row_nnz, block_n=None, block_nnz=None,
                                     max_segments=None):
    """block_n / block_nnz defaults and the explicit max_segments check
    (spmm_csr.py:299-334)."""
    warp_size, factor = _select_spmm_alg1_warp_and_factor(n_dense_cols)
    if block_n is None:
        block_n = warp_size * factor
    if block_nnz is None:
        block_nnz = warp_size
    if block_n <= 0 or block_nnz <= 0:
        raise ValueError("block_n and block_nnz must be positive when provided")
    if max_segments is not None and max_segments <= 0:
        raise ValueError("max_segments must be positive when provided")
    required = (max_row_nnz + block_nnz - 1) // block_nnz if max_row_nnz > 0 else 0
    if max_segments is not None and required > int(max_segments):
        raise ValueError(
            "row nnz requires more CSR segments than the explicit max_segments override allows: "
            f"required {required}, provided {int(max_segments)}")
    return {"block_n": int(block_n), "block_nnz": int(block_nnz),
            "max_segments": None if max_segments is None else int(max_segments),
            "required_segments": int(required), "warp_size": int(warp_size),
            "factor": int(factor), "max_row_nnz": int(max_row_nnz),
            "auto_max_segments": max_segments is None}


# ---------------------------------------------------------------- launches

def _plan_for(cache, kernel_indptr, n_rows, dtype, accumulate, schedule, row_lengths):
    key = (int(accumulate), schedule.key())
    p = cache.get(key)
    if p is None:
        p = RowSplitPlan(kernel_indptr, n_rows, dtype, schedule, accumulate,
                         row_lengths=row_lengths)
        cache[key] = p
    return p


def _spmm_target(out, beta):
    """Where the kernel writes C: ``out`` itself when it is contiguous; for a
    strided ``out``, a contiguous copy of it when beta != 0 (the kernel reads
    C), else a new tensor; the caller copies the result back into ``out``."""
    if out is None or out.is_contiguous():
        return out
    return out.contiguous() if float(beta) != 0.0 else None


def _launch_spmm(csr, plan, B, n_rows, n_dense, dtype, device, out=None, alpha=1.0, beta=0.0,
                 col_tile=0, panels=1):
    """One launch over all of B's columns, or with panels > 1 one launch per
    column panel (panel width a multiple of 4), panel after panel, so the B
    columns gathered at a time are a 1/panels share of B."""
    C = out if out is not None else torch.empty((n_rows, n_dense), dtype=dtype, device=device)
    if n_rows == 0 or n_dense == 0:
        return C
    ldb = B.stride(0) if B.numel() else n_dense
    width = n_dense
    if panels > 1:
        per = -(-n_dense // panels)
        width = -(-per // 4) * 4
    for c0 in range(0, n_dense, width):
        w = min(width, n_dense - c0)
        Bp = B[:, c0:c0 + w] if c0 or w != n_dense else B
        Cp = C[:, c0:c0 + w] if c0 or w != n_dense else C
        _lib.spmm(csr, Bp, ldb, Cp, C.stride(0), w, alpha, beta, plan.struct(w),
                  min(col_tile, w) if col_tile > 0 else 0, device)
    return C


def _hip_spmm_csr_impl(data, indices, indptr, B, n_rows, n_dense_cols, block_n, block_nnz,
                       schedule=None, row_lengths=None, n_cols=None, out=None, alpha=1.0,
                       beta=0.0):
    """ALG1 default path (spmm_csr.py:525-599): fp32 widened to fp64 in the
    kernel, fp16 / bf16 in fp32, fp64 / complex natively."""
    device, dtype = data.device, data.dtype
    if n_rows == 0 or n_dense_cols == 0 or B.shape[0] == 0:
        if out is None:
            return torch.zeros((n_rows, n_dense_cols), dtype=dtype, device=device)
        return out.mul_(beta) if beta != 0.0 else out.zero_()
    n_cols = int(B.shape[0]) if n_cols is None else n_cols
    acc = _lib.FS_ACC_WIDE if dtype == torch.float32 else _lib.FS_ACC_NATIVE
    sch = schedule or default_spmm_schedule(n_rows, data.numel(), n_dense_cols)
    plan = RowSplitPlan(indptr, n_rows, dtype, sch, acc, row_lengths=row_lengths)
    csr = _lib.make_csr(data, indices, indptr, n_rows, n_cols)
    return _launch_spmm(csr, plan, B, n_rows, n_dense_cols, dtype, device, out, alpha, beta,
                        sch.col_tile)


_triton_spmm_csr_impl = _hip_spmm_csr_impl


def flagsparse_spmm_csr(data, indices, indptr, B, shape, block_n=None, block_nnz=None,
                        max_segments=None, out=None, return_time=False, alpha=1.0, beta=0.0):
    """C = A @ B (spmm_csr.py:2088-2157): row-major, non-transposed B and C.

    ``block_n`` / ``block_nnz`` / ``max_segments`` keep the reference's checks;
    they do not change the HIP schedule.  Every call validates the matrix
    with one device pass whose results are read on the host (one device
    synchronisation, where the reference makes two); repeated products with
    one matrix are cheaper through ``prepare_spmm_csr_opt``.
    """
    if block_n is not None and block_n <= 0:
        raise ValueError("block_n must be positive when provided")
    if block_nnz is not None and block_nnz <= 0:
        raise ValueError("block_nnz must be positive when provided")
    if max_segments is not None and max_segments <= 0:
        raise ValueError("max_segments must be positive when provided")
    (data, kernel_indices, kernel_indptr, B, n_rows, n_cols, n_dense_cols, row_lengths,
     max_row_nnz) = _prepare_spmm_csr_inputs(data, indices, indptr, B, shape)
    _resolve_spmm_alg1_launch_config(n_dense_cols, max_row_nnz, block_n=block_n,
                                     block_nnz=block_nnz, max_segments=max_segments)
    if out is not None:
        if not out.is_cuda:
            raise ValueError("out must be a CUDA tensor")
        if out.device != data.device:
            raise ValueError("out must be on the same CUDA device as the inputs")
        if tuple(out.shape) != (n_rows, n_dense_cols) or out.dtype != data.dtype:
            raise ValueError("out shape/dtype must match result")
    if float(beta) != 0.0 and out is None:
        raise ValueError("beta != 0 requires out (the C that is scaled and accumulated)")
    target = _spmm_target(out, beta)
    t0 = None
    if return_time:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
    C = _hip_spmm_csr_impl(data, kernel_indices, kernel_indptr, B, n_rows, n_dense_cols,
                           block_n=block_n, block_nnz=block_nnz, row_lengths=row_lengths,
                           n_cols=n_cols, out=target, alpha=alpha, beta=beta)
    elapsed_ms = None
    if return_time:
        torch.cuda.synchronize()
        elapsed_ms = (time.perf_counter() - t0) * 1000.0
    if out is not None and C is not out:
        out.copy_(C)
        C = out
    if return_time:
        return C, elapsed_ms
    return C


# --------------------------------------------------------------------- opt

class PreparedCsrSpmmOpt:
    """Cached metadata for repeated SpMM-opt calls (spmm_csr.py:337-392),
    including the reference's bucket lists and 256-entry split parts (for
    introspection; the HIP kernel runs the row-split plan)."""

    __slots__ = (
        "data", "kernel_indices", "kernel_indptr", "shape", "n_rows", "n_cols", "row_lengths",
        "max_row_nnz", "row_buckets", "supports_opt", "long_part_rows", "long_part_starts",
        "long_part_ends", "long_row_ids", "long_row_part_ptr", "long_row_fallback_only",
        "schedule", "library_schedule", "_plans", "_csr", "schedule_record",
    )

    def __init__(self, data, kernel_indices, kernel_indptr, shape, n_rows, n_cols, row_lengths,
                 max_row_nnz, row_buckets, long_part_rows, long_part_starts, long_part_ends,
                 long_row_ids, long_row_part_ptr, long_row_fallback_only, schedule=None,
                 library_schedule=False):
        self.data = data
        self.kernel_indices = kernel_indices
        self.kernel_indptr = kernel_indptr
        self.shape = (int(shape[0]), int(shape[1]))
        self.n_rows = int(n_rows)
        self.n_cols = int(n_cols)
        self.row_lengths = row_lengths
        self.max_row_nnz = int(max_row_nnz)
        self.row_buckets = row_buckets
        self.supports_opt = data.dtype in (torch.float32, torch.float64)
        self.long_part_rows = long_part_rows
        self.long_part_starts = long_part_starts
        self.long_part_ends = long_part_ends
        self.long_row_ids = long_row_ids
        self.long_row_part_ptr = long_row_part_ptr
        self.long_row_fallback_only = bool(long_row_fallback_only)
        self.schedule = schedule
        self.library_schedule = bool(library_schedule)  # chosen by the library, not the caller
        self.schedule_record = None  # how the library chose `schedule`
        self._plans = {}
        self._csr = None

    def csr_struct(self):
        if self._csr is None:
            self._csr = _lib.make_csr(self.data, self.kernel_indices, self.kernel_indptr,
                                      self.n_rows, self.n_cols)
        return self._csr

    def plan(self, accumulate, n_dense, schedule=None):
        sch = schedule or self.schedule or default_spmm_schedule(self.n_rows, self.data.numel(),
                                                                 n_dense)
        return _plan_for(self._plans, self.kernel_indptr, self.n_rows, self.data.dtype,
                         accumulate, sch, self.row_lengths)


_SPMM_OPT_BUCKET_SPECS = (
    {"max_row_nnz": 32, "kind": "batched", "batch_rows": 8, "block_nnz": 32},
    {"max_row_nnz": 128, "kind": "batched", "batch_rows": 4, "block_nnz": 64},
    {"max_row_nnz": 512, "kind": "vector", "batch_rows": 1, "block_nnz": 128},
    {"max_row_nnz": 2048, "kind": "vector", "batch_rows": 1, "block_nnz": 128},
    {"max_row_nnz": None, "kind": "split", "batch_rows": 1, "block_nnz": 256},
)
_SPMM_OPT_LONG_ROW_THRESHOLD = 2048


def _select_spmm_opt_block_n(n_dense_cols):
    """Column tile of the opt path (spmm_csr.py:405-414)."""
    for cap in (8, 16, 32, 64):
        if n_dense_cols <= cap:
            return cap
    return 128


def _build_spmm_opt_split_metadata(kernel_indptr, long_rows, part_block_nnz):
    """Parts [s, min(s + part, e)) of each long row (spmm_csr.py:417-441),
    computed on the device: (part_rows, part_starts, part_ends, row_part_ptr)."""
    device = kernel_indptr.device
    rows64 = long_rows.to(torch.int64)
    starts = kernel_indptr[rows64].to(torch.int64)
    ends = kernel_indptr[rows64 + 1].to(torch.int64)
    per = (ends - starts + part_block_nnz - 1) // part_block_nnz
    row_part_ptr = torch.zeros(rows64.numel() + 1, dtype=torch.int64, device=device)
    if rows64.numel():
        torch.cumsum(per, 0, out=row_part_ptr[1:])
    owner = torch.repeat_interleave(torch.arange(rows64.numel(), device=device), per)
    k = torch.arange(owner.numel(), device=device, dtype=torch.int64) - row_part_ptr[owner]
    p_start = starts[owner] + k * part_block_nnz
    p_end = torch.minimum(p_start + part_block_nnz, ends[owner])
    row_dtype = long_rows.dtype if long_rows.numel() > 0 else torch.int32
    return rows64[owner].to(row_dtype), p_start, p_end, row_part_ptr


def _build_spmm_opt_buckets(row_lengths):
    """Row lists of the five opt buckets (spmm_csr.py:444-477)."""
    device = row_lengths.device
    row_index_dtype = torch.int32 if int(row_lengths.numel()) - 1 <= _INDEX_LIMIT_INT32 else torch.int64
    buckets = []
    lower = 0
    long_rows = torch.empty((0,), dtype=row_index_dtype, device=device)
    for spec in _SPMM_OPT_BUCKET_SPECS:
        upper = spec["max_row_nnz"]
        if upper is None:
            mask = row_lengths > lower
        elif lower == 0:
            mask = row_lengths <= upper
        else:
            mask = (row_lengths > lower) & (row_lengths <= upper)
        rows = torch.nonzero(mask, as_tuple=False).flatten().to(row_index_dtype)
        if rows.numel() > 0:
            buckets.append({"kind": spec["kind"], "rows": rows, "batch_rows": spec["batch_rows"],
                            "block_nnz": spec["block_nnz"]})
            if spec["kind"] == "split":
                long_rows = rows
        if upper is not None:
            lower = upper
    return buckets, long_rows


def prepare_spmm_csr_opt(data, indices, indptr, shape, schedule=None):
    """Validate once and cache metadata (spmm_csr.py:480-522).  ``schedule``:
    a ``Schedule``, "learned" (default when a model is installed) or
    "default"."""
    stats_out = []
    (data, kernel_indices, kernel_indptr, n_rows, n_cols, row_lengths,
     max_row_nnz) = _prepare_spmm_csr_matrix(data, indices, indptr, shape, _stats_out=stats_out)
    library = not isinstance(schedule, Schedule)
    schedule, how = choose_schedule("spmm64", schedule, data, n_rows, n_cols, stats_out[0])
    row_buckets, long_rows = _build_spmm_opt_buckets(row_lengths)
    part_rows, part_starts, part_ends, part_ptr = _build_spmm_opt_split_metadata(
        kernel_indptr, long_rows, 256)
    prepared = PreparedCsrSpmmOpt(
        data=data, kernel_indices=kernel_indices, kernel_indptr=kernel_indptr, shape=shape,
        n_rows=n_rows, n_cols=n_cols, row_lengths=row_lengths, max_row_nnz=max_row_nnz,
        row_buckets=row_buckets, long_part_rows=part_rows, long_part_starts=part_starts,
        long_part_ends=part_ends, long_row_ids=long_rows, long_row_part_ptr=part_ptr,
        long_row_fallback_only=False, schedule=schedule, library_schedule=library)
    prepared.schedule_record = {"schedule": schedule, "chosen_by": how}
    return prepared


def _check_opt_B(prepared, B):
    if not prepared.supports_opt:
        raise TypeError("spmm opt only supports float32 and float64")
    if not B.is_cuda:
        raise ValueError("B must be a CUDA tensor")
    if B.device != prepared.data.device:
        raise ValueError("B must be on the same CUDA device as the sparse matrix")
    if B.dtype != prepared.data.dtype:
        raise TypeError("B dtype must match sparse matrix dtype")
    if B.shape[0] != prepared.n_cols:
        raise ValueError(f"B.shape[0] must be n_cols={prepared.n_cols}, got {B.shape[0]}")
    if B.ndim != 2:
        raise ValueError("B must be a 2D dense tensor")


def _library_spmm_launch(sch, dtype, n_dense):
    """(schedule, column tile) to launch for a schedule the library chose
    itself (selector, fixed rule or exhaustive search).  Those are picked for
    B with 64 columns at 4 entries in flight per lane; measured on MI355X
    (1M-row banded matrix, N = 32 / 64 / 128), fp32 runs faster with 8 in
    flight (418 vs 437 us at N = 64), and a wider B keeps the chosen lanes per
    64 columns (N = 128 with one 128-column tile: 883 vs 926 us).  fp64 keeps
    4 in flight (851 vs 929 us).  Schedules passed in by the caller run as
    given."""
    if dtype == torch.float32 and sch.unroll == 4:
        sch = dataclasses.replace(sch, unroll=8)
    col_tile = sch.col_tile
    if col_tile > 0 and n_dense > 64:
        col_tile *= n_dense // 64
    return sch, col_tile


def spmm_launch_schedule(prepared, n_dense, schedule=None):
    """The (schedule, column tile) an opt / alg2 SpMM call with n_dense
    columns launches."""
    if schedule is None and prepared.library_schedule:
        sch = prepared.schedule or default_spmm_schedule(prepared.n_rows, prepared.data.numel(),
                                                         n_dense)
        return _library_spmm_launch(sch, prepared.data.dtype, n_dense)
    sch = schedule or prepared.schedule or default_spmm_schedule(prepared.n_rows,
                                                                 prepared.data.numel(), n_dense)
    return sch, sch.col_tile


def _hip_spmm_csr_impl_opt_prepared(prepared, B, out=None, alpha=1.0, beta=0.0, schedule=None):
    """Opt path (spmm_csr.py:2034-2056): native accumulation; returns (C, False)."""
    _check_opt_B(prepared, B)
    B = B.contiguous()
    n_dense = int(B.shape[1])
    sch, col_tile = spmm_launch_schedule(prepared, n_dense, schedule)
    plan = prepared.plan(_lib.FS_ACC_NATIVE, n_dense, sch)
    C = _launch_spmm(prepared.csr_struct(), plan, B, prepared.n_rows, n_dense, B.dtype, B.device,
                     out, alpha, beta, col_tile, sch.b_panels)
    return C, False


_triton_spmm_csr_impl_opt_prepared = _hip_spmm_csr_impl_opt_prepared


# ------------------------------------------------------- diagnose candidate
#
# The reference's diagnose script (tests/diagnose_spmm_opt.py:311-330, 527)
# labels rows by the candidate bucket table of spmm_csr.py:1457-1497 and
# compares a candidate SpMM against the opt path.  The table is kept as
# returned metadata; the candidate values come from the HIP SpMM.

_SPMM_OPT_CANDIDATE_BUCKET_SPECS_F32 = (
    {"label": "short_16", "kind": "micro", "min_row_nnz": 0, "max_row_nnz": 16, "block_nnz": 16},
    {"label": "short_32", "kind": "micro", "min_row_nnz": 17, "max_row_nnz": 32, "block_nnz": 32},
    {"label": "short_64", "kind": "segmented", "min_row_nnz": 33, "max_row_nnz": 64,
     "block_nnz": 64, "segments": 2},
    {"label": "short_128", "kind": "segmented", "min_row_nnz": 65, "max_row_nnz": 128,
     "block_nnz": 128, "segments": 4},
    {"label": "vector_192", "kind": "segmented", "min_row_nnz": 129, "max_row_nnz": 192,
     "block_nnz": 64, "segments": 2},
    {"label": "vector_256", "kind": "segmented", "min_row_nnz": 193, "max_row_nnz": 256,
     "block_nnz": 64, "segments": 4},
    {"label": "vector_512", "kind": "segmented", "min_row_nnz": 257, "max_row_nnz": 512,
     "block_nnz": 64, "segments": 4},
    {"label": "vector_1024", "kind": "segmented", "min_row_nnz": 513, "max_row_nnz": 1024,
     "block_nnz": 32, "segments": 8},
    {"label": "vector_2048", "kind": "segmented", "min_row_nnz": 1025, "max_row_nnz": 2048,
     "block_nnz": 32, "segments": 8},
)
_SPMM_OPT_CANDIDATE_BUCKET_SPECS_F64 = (
    {"label": "short_16", "kind": "segmented", "min_row_nnz": 0, "max_row_nnz": 16,
     "block_nnz": 16, "segments": 2},
    {"label": "short_32", "kind": "segmented", "min_row_nnz": 17, "max_row_nnz": 32,
     "block_nnz": 32, "segments": 2},
    {"label": "short_64", "kind": "segmented", "min_row_nnz": 33, "max_row_nnz": 64,
     "block_nnz": 64, "segments": 2},
    {"label": "short_128", "kind": "segmented", "min_row_nnz": 65, "max_row_nnz": 128,
     "block_nnz": 128, "segments": 4},
    {"label": "vector_256", "kind": "segmented", "min_row_nnz": 129, "max_row_nnz": 256,
     "block_nnz": 64, "segments": 4},
    {"label": "vector_512", "kind": "segmented", "min_row_nnz": 257, "max_row_nnz": 512,
     "block_nnz": 64, "segments": 4},
    {"label": "vector_1024", "kind": "segmented", "min_row_nnz": 513, "max_row_nnz": 1024,
     "block_nnz": 32, "segments": 8},
    {"label": "vector_2048", "kind": "segmented", "min_row_nnz": 1025, "max_row_nnz": 2048,
     "block_nnz": 32, "segments": 8},
)
_SPMM_OPT_CANDIDATE_SPLIT_THRESHOLD_F32 = 2048
_SPMM_OPT_CANDIDATE_SPLIT_THRESHOLD_F64 = 2048


def _select_spmm_opt_candidate_specs(dtype):
    return (_SPMM_OPT_CANDIDATE_BUCKET_SPECS_F64 if dtype == torch.float64
            else _SPMM_OPT_CANDIDATE_BUCKET_SPECS_F32)


def _select_spmm_opt_candidate_split_threshold(dtype):
    return (_SPMM_OPT_CANDIDATE_SPLIT_THRESHOLD_F64 if dtype == torch.float64
            else _SPMM_OPT_CANDIDATE_SPLIT_THRESHOLD_F32)


def _build_spmm_opt_candidate_buckets(prepared):
    """Rows of each candidate bucket (spmm_csr.py:1835-1873): a length range
    per spec (the first takes every row up to its maximum, empty rows
    included), then a "split" bucket of the rows over the threshold.  Row
    ids are int32 unless the row count needs int64."""
    row_lengths = prepared.row_lengths
    device = row_lengths.device
    row_count = int(row_lengths.numel())
    index_dtype = torch.int32 if row_count - 1 <= _INDEX_LIMIT_INT32 else torch.int64
    all_rows = torch.arange(row_count, device=device, dtype=index_dtype)
    buckets = []
    for spec in _select_spmm_opt_candidate_specs(prepared.data.dtype):
        lo, hi = int(spec["min_row_nnz"]), int(spec["max_row_nnz"])
        mask = row_lengths <= hi if lo <= 0 else (row_lengths >= lo) & (row_lengths <= hi)
        rows = all_rows[mask]
        if rows.numel() == 0:
            continue
        bucket = {"label": spec["label"], "kind": spec["kind"], "rows": rows,
                  "block_nnz": int(spec["block_nnz"]), "min_row_nnz": lo, "max_row_nnz": hi}
        if "segments" in spec:
            bucket["segments"] = int(spec["segments"])
        buckets.append(bucket)
    split_rows = all_rows[row_lengths > _select_spmm_opt_candidate_split_threshold(
        prepared.data.dtype)]
    if split_rows.numel() > 0:
        buckets.append({"label": "split", "kind": "split", "rows": split_rows, "block_nnz": 256})
    return buckets


def _hip_spmm_csr_impl_opt_prepared_candidate(prepared, B):
    """The diagnose candidate (spmm_csr.py:1995-2026): opt-path validation,
    then C with every row accumulated in fp64 and rounded once (the
    reference's candidate compensates short rows with Kahan sums; fp64
    accumulation is at least as accurate).  Returns (C, False): no long-row
    fallback exists here."""
    _check_opt_B(prepared, B)
    B = B.contiguous()
    n_dense = int(B.shape[1])
    sch, col_tile = spmm_launch_schedule(prepared, n_dense)
    plan = prepared.plan(_lib.FS_ACC_WIDE, n_dense, sch)
    C = _launch_spmm(prepared.csr_struct(), plan, B, prepared.n_rows, n_dense, B.dtype, B.device,
                     None, 1.0, 0.0, col_tile)
    return C, False


_triton_spmm_csr_impl_opt_prepared_candidate = _hip_spmm_csr_impl_opt_prepared_candidate


def _flagsparse_spmm_csr_opt_candidate_for_diagnose(prepared, B):
    """C of the diagnose candidate (spmm_csr.py:2029-2031)."""
    C_out, _ = _hip_spmm_csr_impl_opt_prepared_candidate(prepared, B)
    return C_out


def flagsparse_spmm_csr_opt(data=None, indices=None, indptr=None, B=None, shape=None,
                            prepared=None, out=None, return_time=False, alpha=1.0, beta=0.0):
    """Native fp32 / fp64 SpMM-opt (spmm_csr.py:2160-2214)."""
    if prepared is not None and not isinstance(prepared, PreparedCsrSpmmOpt):
        raise TypeError("prepared must be a PreparedCsrSpmmOpt instance")
    if prepared is None:
        if any(a is None for a in (data, indices, indptr, shape)):
            raise ValueError(
                "data, indices, indptr, and shape are required when prepared is not provided")
        prepared = prepare_spmm_csr_opt(data, indices, indptr, shape)
    elif shape is not None:
        sh = (int(shape[0]), int(shape[1]))
        if sh != prepared.shape:
            raise ValueError(f"shape {sh} does not match prepared.shape {prepared.shape}")
    if B is None:
        raise ValueError("B is required")
    if not prepared.supports_opt:
        raise TypeError("flagsparse_spmm_csr_opt only supports float32 and float64")
    if not B.is_cuda:
        raise ValueError("B must be a CUDA tensor")
    if B.device != prepared.data.device:
        raise ValueError("B must be on the same CUDA device as sparse matrix data")
    if out is not None:
        if not out.is_cuda:
            raise ValueError("out must be a CUDA tensor")
        if out.device != prepared.data.device:
            raise ValueError("out must be on the same CUDA device as sparse matrix data")
        if tuple(out.shape) != (prepared.n_rows, int(B.shape[1])) or out.dtype != prepared.data.dtype:
            raise ValueError("out shape/dtype must match result")
    if float(beta) != 0.0 and out is None:
        raise ValueError("beta != 0 requires out (the C that is scaled and accumulated)")
    target = _spmm_target(out, beta)
    t0 = None
    if return_time:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
    C, _ = _triton_spmm_csr_impl_opt_prepared(prepared, B, out=target, alpha=alpha, beta=beta)
    elapsed_ms = None
    if return_time:
        torch.cuda.synchronize()
        elapsed_ms = (time.perf_counter() - t0) * 1000.0
    if out is not None and C is not out:
        out.copy_(C)
        C = out
    if return_time:
        return C, elapsed_ms
    return C


def _spmm_opt_reference_error(candidate, reference, value_dtype):
    """max(|d| / (atol + rtol |ref|)) with the opt tolerances (spmm_csr.py:2217-2223)."""
    atol, rtol = _spmm_coo_reference_tolerance(value_dtype)
    if candidate.numel() == 0:
        return 0.0
    return float(torch.max(torch.abs(candidate - reference) /
                           (atol + rtol * torch.abs(reference))).item())


def _torch_reference_spmm(data, indices, indptr, B, shape, value_dtype):
    """torch.sparse.mm in the promoted dtype (fp32 -> fp64), as the reference's
    benchmarks do (spmm_csr.py:2265-2278)."""
    ref_dtype = torch.float64 if value_dtype == torch.float32 else value_dtype
    csr = torch.sparse_csr_tensor(indptr.to(torch.int64), indices.to(torch.int64),
                                  data.to(ref_dtype), size=shape, device=data.device)
    return torch.sparse.mm(csr, B.to(ref_dtype)).to(value_dtype)


def benchmark_spmm_opt_case(n_rows=4096, n_cols=4096, nnz=65536, n_dense_cols=32,
                            value_dtype=torch.float32, index_dtype=torch.int32, warmup=20,
                            iters=200, run_cusparse=True):
    """Base vs opt timing and error against the promoted torch reference
    (spmm_csr.py:2226-2348).  ``cu_ms`` times the vendor library, rocSPARSE."""
    if value_dtype not in (torch.float32, torch.float64):
        raise TypeError("benchmark_spmm_opt_case only supports float32 and float64")
    device = torch.device("cuda")
    data, indices, indptr = _build_random_csr(n_rows, n_cols, nnz, value_dtype, index_dtype, device)
    B = _build_random_dense((n_cols, n_dense_cols), value_dtype, device)
    shape = (n_rows, n_cols)
    prepared = prepare_spmm_csr_opt(data, indices, indptr, shape)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    base = flagsparse_spmm_csr(data, indices, indptr, B, shape)
    torch.cuda.synchronize()
    base_first = (time.perf_counter() - t0) * 1000.0
    base, base_ms = _benchmark_cuda_op(lambda: flagsparse_spmm_csr(data, indices, indptr, B, shape),
                                       warmup, iters)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    opt = flagsparse_spmm_csr_opt(B=B, prepared=prepared)
    torch.cuda.synchronize()
    opt_first = (time.perf_counter() - t0) * 1000.0
    opt, opt_ms = _benchmark_cuda_op(lambda: flagsparse_spmm_csr_opt(B=B, prepared=prepared),
                                     warmup, iters)
    ref = _torch_reference_spmm(data, indices, indptr, B, shape, value_dtype)
    pt_ms = None
    try:
        pt = torch.sparse_csr_tensor(indptr.to(torch.int64), indices.to(torch.int64), data,
                                     size=shape, device=device)
        _, pt_ms = _benchmark_cuda_op(lambda: torch.sparse.mm(pt, B), warmup, iters)
    except Exception:
        pt_ms = None
    err_base = _spmm_opt_reference_error(base, ref, value_dtype)
    err_opt = _spmm_opt_reference_error(opt, ref, value_dtype)
    cu_ms = None
    if run_cusparse:
        from .. import vendor
        _, run, _reason = vendor.spmm_baseline(data, indices, indptr, shape, B)
        if run is not None:
            _, cu_ms = _benchmark_cuda_op(run, warmup, iters)
    return {
        "parameters": {"n_rows": n_rows, "n_cols": n_cols, "nnz": nnz,
                       "n_dense_cols": n_dense_cols, "value_dtype": str(value_dtype),
                       "index_dtype": str(index_dtype)},
        "performance": {"base_ms": base_ms, "base_first_call_ms": base_first, "opt_ms": opt_ms,
                        "opt_first_call_ms": opt_first, "pt_ms": pt_ms, "cu_ms": cu_ms,
                        "opt_vs_base": base_ms / opt_ms if opt_ms > 0 else None,
                        "opt_vs_pt": pt_ms / opt_ms if pt_ms is not None and opt_ms > 0 else None,
                        "opt_vs_cu": cu_ms / opt_ms if cu_ms is not None and opt_ms > 0 else None},
        "verification": {"err_base": err_base, "err_opt": err_opt, "base_ok": err_base <= 1.0,
                         "opt_ok": err_opt <= 1.0, "status": "PASS" if err_opt <= 1.0 else "FAIL"},
        "backend_status": {"long_row_fallback_used": bool(prepared.long_row_fallback_only),
                           "flagsparse_internal_route": "hip-row-split"},
        "samples": {"base": base, "opt": opt, "reference": ref},
    }


__all__ = [
    "SUPPORTED_SPMM_VALUE_DTYPES",
    "PreparedCsrSpmmOpt",
    "flagsparse_spmm_csr",
    "flagsparse_spmm_csr_opt",
    "prepare_spmm_csr_opt",
    "benchmark_spmm_case",
    "benchmark_spmm_opt_case",
    "comprehensive_spmm_test",
    "Schedule",
]


def benchmark_spmm_case(n_rows=4096, n_cols=4096, nnz=65536, n_dense_cols=32,
                        value_dtype=torch.float32, index_dtype=torch.int32, warmup=20, iters=200,
                        block_n=None, block_nnz=None, max_segments=None, run_cusparse=True):
    """ALG1 SpMM vs torch.sparse.mm and the vendor library on one random matrix;
    parameters and result layout of spmm_csr.py:2351-2564.  The vendor
    library is rocSPARSE (its numbers fill the ``cusparse_*`` fields)."""
    from .. import vendor
    device = torch.device("cuda")
    data, indices, indptr = _build_random_csr(n_rows, n_cols, nnz, value_dtype, index_dtype, device)
    B = _build_random_dense((n_cols, n_dense_cols), value_dtype, device)
    shape = (n_rows, n_cols)
    max_row_nnz = int(torch.max(indptr[1:] - indptr[:-1]).item()) if n_rows > 0 else 0
    launch = _resolve_spmm_alg1_launch_config(n_dense_cols, max_row_nnz, block_n=block_n,
                                              block_nnz=block_nnz, max_segments=max_segments)
    kwargs = dict(data=data, indices=indices, indptr=indptr, B=B, shape=shape,
                  block_n=launch["block_n"], block_nnz=launch["block_nnz"],
                  max_segments=launch["max_segments"], return_time=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    flagsparse_spmm_csr(**kwargs)
    torch.cuda.synchronize()
    first_ms = (time.perf_counter() - t0) * 1000.0
    hip_C, hip_ms = _benchmark_cuda_op(lambda: flagsparse_spmm_csr(**kwargs), warmup, iters)

    indptr64, indices64 = indptr.to(torch.int64), indices.to(torch.int64)
    pytorch_reason, pytorch_ms, pytorch_format = None, None, "CSR"
    promote = {torch.float16: torch.float32, torch.bfloat16: torch.float32,
               torch.float32: torch.float64, torch.complex64: torch.complex128}
    ref_dtype = promote.get(value_dtype, value_dtype)
    try:
        csr_pt = torch.sparse_csr_tensor(indptr64, indices64, data, size=shape, device=device)
        pytorch_op = lambda: torch.sparse.mm(csr_pt, B)  # noqa: E731
        csr_ref = torch.sparse_csr_tensor(indptr64, indices64, data.to(ref_dtype), size=shape,
                                          device=device)
        expected = torch.sparse.mm(csr_ref, B.to(ref_dtype)).to(value_dtype)
    except Exception as exc:
        pytorch_format = "COO"
        pytorch_reason = f"CSR fallback: {exc}"
        rows = torch.repeat_interleave(torch.arange(n_rows, device=device, dtype=torch.int64),
                                       indptr64[1:] - indptr64[:-1])
        coo = torch.sparse_coo_tensor(torch.stack([rows, indices64]), data, shape,
                                      device=device).coalesce()
        pytorch_op = lambda: torch.sparse.mm(coo, B)  # noqa: E731
        expected = torch.sparse.mm(coo.to(ref_dtype), B.to(ref_dtype)).to(value_dtype)
    pytorch_values = expected
    try:
        pytorch_values, pytorch_ms = _benchmark_cuda_op(pytorch_op, warmup, iters)
    except Exception as exc:
        pytorch_reason = str(exc) if pytorch_reason is None else f"{pytorch_reason}; timing: {exc}"
    hip_metrics = _spmm_validation_metrics(hip_C, expected)

    vendor_ms = vendor_values = vendor_metrics = vendor_match = None
    vendor_reason = None
    if run_cusparse:
        _, run, vendor_reason = vendor.spmm_baseline(data, indices, indptr, shape, B)
        if run is not None:
            try:
                vendor_values, vendor_ms = _benchmark_cuda_op(run, warmup, iters)
                vendor_values = vendor_values.clone()
                vendor_metrics = _spmm_validation_metrics(vendor_values, expected)
                vendor_match = vendor_metrics["strict_allclose_match"]
            except Exception as exc:
                vendor_reason = str(exc)
    threshold = _spmm_relative_threshold(value_dtype)

    def vm(key):
        return vendor_metrics[key] if vendor_metrics is not None else None

    return {
        "parameters": {
            "format": "csr", "n_rows": n_rows, "n_cols": n_cols, "nnz": nnz,
            "n_dense_cols": n_dense_cols, "value_dtype": str(value_dtype),
            "index_dtype": str(index_dtype), "warmup": warmup, "iters": iters,
            "block_n": launch["block_n"], "block_nnz": launch["block_nnz"],
            "max_segments": launch["max_segments"],
            "required_segments": launch["required_segments"],
            "alg1_warp_size": launch["warp_size"], "alg1_factor": launch["factor"],
            "auto_max_segments": launch["auto_max_segments"], "run_cusparse": run_cusparse,
        },
        "performance": {
            "pytorch_ms": pytorch_ms,
            "triton_ms": hip_ms,
            "triton_first_call_ms": first_ms,
            "cusparse_ms": vendor_ms,
            "triton_speedup_vs_pytorch": (pytorch_ms / hip_ms
                                          if pytorch_ms is not None and hip_ms > 0 else None),
            "triton_speedup_vs_cusparse": (vendor_ms / hip_ms
                                           if vendor_ms is not None and hip_ms > 0 else None),
        },
        "verification": {
            "triton_match_reference": hip_metrics["strict_allclose_match"],
            "triton_match_pytorch": hip_metrics["strict_allclose_match"],
            "triton_max_error": hip_metrics["max_abs_error"],
            "triton_max_abs_error": hip_metrics["max_abs_error"],
            "triton_max_relative_error": hip_metrics["max_relative_error"],
            "triton_sum_relative_error": hip_metrics["sum_relative_error"],
            "triton_relative_threshold": hip_metrics["relative_threshold"],
            "triton_strict_allclose_match": hip_metrics["strict_allclose_match"],
            "pytorch_match_reference": True,
            "pytorch_max_error": 0.0,
            "pytorch_max_abs_error": 0.0,
            "pytorch_max_relative_error": 0.0,
            "pytorch_sum_relative_error": 0.0,
            "pytorch_relative_threshold": threshold,
            "cusparse_match_reference": vendor_match,
            "cusparse_match_pytorch": vendor_match,
            "cusparse_max_error": vm("max_abs_error"),
            "cusparse_max_abs_error": vm("max_abs_error"),
            "cusparse_max_relative_error": vm("max_relative_error"),
            "cusparse_sum_relative_error": vm("sum_relative_error"),
            "cusparse_relative_threshold": (vendor_metrics["relative_threshold"]
                                            if vendor_metrics is not None else threshold),
            "cusparse_strict_allclose_match": vm("strict_allclose_match"),
        },
        "backend_status": {
            "pytorch_unavailable_reason": pytorch_reason,
            "pytorch_sparse_format": pytorch_format,
            "cusparse_unavailable_reason": vendor_reason,
            "vendor_library": "rocSPARSE",
            "flagsparse_internal_route": "csr-alg1",
        },
        "samples": {"pytorch": pytorch_values, "triton": hip_C, "reference": expected,
                    "cusparse": vendor_values},
    }


def comprehensive_spmm_test(n_rows=4096, n_cols=4096, nnz=65536, n_dense_cols=32,
                            dtype=torch.float32, index_dtype=torch.int32, warmup=20, iters=200,
                            block_n=None, block_nnz=None, max_segments=None, run_cusparse=True):
    """One-configuration SpMM benchmark entry (spmm_csr.py:2565-2593)."""
    return benchmark_spmm_case(n_rows=n_rows, n_cols=n_cols, nnz=nnz, n_dense_cols=n_dense_cols,
                               value_dtype=dtype, index_dtype=index_dtype, warmup=warmup,
                               iters=iters, block_n=block_n, block_nnz=block_nnz,
                               max_segments=max_segments, run_cusparse=run_cusparse)
