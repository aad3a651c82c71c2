"""Split-precision value storage for fp32 CSR SpMV.

``prepare_spmv_csr_split`` validates an fp32 CSR matrix like
``prepare_spmv_csr`` and re-encodes its values on the device into one of:

* ``"bf16x2"``        -- hi = bf16(v), lo = bf16(v - hi); value = hi + lo
                         (about 16 significant bits; 4 bytes / entry);
* ``"bf16"``          -- the hi plane only (8 significant bits; 2 bytes);
* ``"f16_rowscale"``  -- fp16(v / s_r) with s_r the power of two at or above
                         the row's max |v| (11 bits; 2 bytes + 4 per row).

``flagsparse_spmv_csr_split`` then computes y = alpha A x + beta y with fp32
x / y and fp32 products, and accumulates in fp32 only: every row sum is a
compensated (TwoSum) fp32 sum -- lane accumulators, the tile roles' row sums
and the carries of rows cut by tile boundaries, chunk partials of long rows,
and the column-sliced combine -- so no fp64 arithmetic is used (fp64 is kept
for float64 calls and the oracle check).  The default schedule is the one
the plain fp32 SpMV of the matrix would use (the packaged selector, else the
fixed rule): for x far beyond the L2 that
is the x-window layout (xwin.py: x gathered from LDS-staged windows, then the
planes streamed in CSR order) or the column-sliced layout (slicing.py, planes
permuted into slice-major order).  The reference stores values only in
the call's dtype; these modes have no counterpart there and are measured
beside the plain fp32 kernel.
"""

from __future__ import annotations

import time

import torch

from . import _lib
from .rowsplit import RowSplitPlan, Schedule
from .schedule_choice import choose_schedule
from .slicing import SlicedCsr
from .sparse_operations.spmv_csr import _prepare_spmv_csr_matrix, _validate_spmv_x
from .xwin import XWinCsr

FORMATS = {
    "bf16x2": _lib.FS_STORE_BF16X2,
    "bf16": _lib.FS_STORE_BF16,
    "f16_rowscale": _lib.FS_STORE_F16_ROWSCALE,
}
# value bytes per entry and per row of each format (for the byte model)
FORMAT_BYTES = {"bf16x2": (4, 0), "bf16": (2, 0), "f16_rowscale": (2, 4)}


class PreparedCsrSpmvSplit:
    """An fp32 CSR matrix with values held in a split-precision format."""

    def __init__(self, kernel_indices, kernel_indptr, shape, value_format, hi, lo, row_scale,
                 row_lengths, max_row_nnz, schedule):
        self.kernel_indices = kernel_indices
        self.kernel_indptr = kernel_indptr
        self.shape = (int(shape[0]), int(shape[1]))
        self.n_rows, self.n_cols = self.shape
        self.value_format = value_format
        self.hi, self.lo, self.row_scale = hi, lo, row_scale
        self.row_lengths = row_lengths
        self.max_row_nnz = max_row_nnz
        self.schedule = schedule
        self.data = torch.empty(0, dtype=torch.float32, device=kernel_indptr.device)  # x / y dtype
        self._plans = {}
        nnz = kernel_indices.numel()
        self._csr = _lib.FsCsr(value_type=_lib.FS_F32, index_type=_lib.FS_I32,
                               ptr_type=_lib.index_code(kernel_indptr.dtype), reserved=0,
                               n_rows=self.n_rows, n_cols=self.n_cols, nnz=nnz,
                               row_ptr=_lib.ptr(kernel_indptr), col_idx=_lib.ptr(kernel_indices),
                               values=None)

    def plan(self, schedule=None):
        sch = schedule or self.schedule
        p = self._plans.get(sch.key())
        if p is None:
            p = RowSplitPlan(self.kernel_indptr, self.n_rows, torch.float32, sch,
                             _lib.FS_ACC_NATIVE, row_lengths=self.row_lengths)
            self._plans[sch.key()] = p
        return p

    def value_bytes(self) -> int:
        per_entry, per_row = FORMAT_BYTES[self.value_format]
        return per_entry * self.kernel_indices.numel() + per_row * self.n_rows


def prepare_spmv_csr_split(data, indices, indptr, shape, value_format="bf16x2", schedule=None):
    """Validate (same checks as prepare_spmv_csr) and encode the values."""
    if value_format not in FORMATS:
        raise ValueError(f"value_format must be one of {sorted(FORMATS)}")
    if torch.is_tensor(data) and data.dtype != torch.float32:
        raise TypeError("split-precision storage is for float32 matrices")
    stats_out = []
    (data, kernel_indices, kernel_indptr, n_rows, n_cols, row_lengths,
     max_row_nnz) = _prepare_spmv_csr_matrix(data, indices, indptr, shape, _stats_out=stats_out)
    if kernel_indices.dtype != torch.int32:
        if n_cols - 1 > 2**31 - 1:
            raise ValueError("split-precision storage needs column indices that fit int32")
        kernel_indices = kernel_indices.to(torch.int32)
    dev = data.device
    nnz = data.numel()
    fmt = FORMATS[value_format]
    hi = torch.empty(nnz, dtype=torch.int16, device=dev)
    lo = torch.empty(nnz, dtype=torch.int16, device=dev) if value_format == "bf16x2" else None
    scale = (torch.empty(n_rows, dtype=torch.float32, device=dev)
             if value_format == "f16_rowscale" else None)
    csr = _lib.make_csr(data, kernel_indices, kernel_indptr, n_rows, n_cols)
    _lib.split_values(fmt, csr, hi, lo, scale, dev)
    if schedule is None:
        # the layout the plain fp32 SpMV of this matrix would use (the packaged
        # selector on the device, else the fixed rule)
        sch, _ = choose_schedule("spmv", None, data, n_rows, n_cols, stats_out[0][0],
                                 kernel_indptr, kernel_indices)
    else:
        sch = schedule
    return PreparedCsrSpmvSplit(kernel_indices, kernel_indptr, (n_rows, n_cols), value_format,
                                hi, lo, scale, row_lengths, max_row_nnz, sch)


def _xwin_split(prepared, sch):
    """The x-window layout of the matrix (columns and row pointers only: the
    split planes are read in CSR order by the second pass)."""
    key = ("xwin", sch.key())
    xw = prepared._plans.get(key)
    if xw is None:
        xw = XWinCsr(None, prepared.kernel_indices, prepared.kernel_indptr, prepared.n_rows,
                     prepared.n_cols, tile_nnz=sch.tile_nnz if sch.tile_nnz > 0 else 2048,
                     window=sch.x_window if sch.x_window > 0 else None,
                     value_dtype=torch.float32)
        prepared._plans[key] = xw
    return xw


def _sliced_split(prepared, sch):
    """Slice-major copy of the encoded planes (see slicing.py); the row
    scales of f16_rowscale follow each compact row's original row, so the
    stored values are exactly those of the unsliced planes."""
    key = ("sliced", sch.key(), _lib.FS_ACC_NATIVE)
    sl = prepared._plans.get(key)
    if sl is None:
        sl = SlicedCsr(None, prepared.kernel_indices, prepared.kernel_indptr, prepared.n_rows,
                       prepared.n_cols, sch.x_slice_cols, sch, _lib.FS_ACC_NATIVE,
                       planes=(prepared.hi, prepared.lo), value_dtype=torch.float32,
                       fp32_sums=True)
        sl.row_scale = (prepared.row_scale[sl.compact_row.to(torch.int64)].contiguous()
                        if prepared.row_scale is not None else None)
        prepared._plans[key] = sl
    return sl


def decoded_values(prepared) -> torch.Tensor:
    """The fp32 values the kernels multiply with (for error reporting)."""
    fmt = prepared.value_format
    hi = prepared.hi
    if fmt == "f16_rowscale":
        h = hi.view(torch.float16).float()
        return h * torch.repeat_interleave(prepared.row_scale, prepared.row_lengths)
    v = (hi.to(torch.int32) << 16).view(torch.float32)
    if fmt == "bf16x2":
        v = v.double() + (prepared.lo.to(torch.int32) << 16).view(torch.float32).double()
    return v.double()


def flagsparse_spmv_csr_split(x, prepared, out=None, alpha=1.0, beta=0.0, return_time=False,
                              schedule=None):
    """y = alpha A x + beta y with split-precision values (see module doc)."""
    x = _validate_spmv_x(x, prepared)
    if out is not None:
        if not out.is_cuda or out.device != x.device:
            raise ValueError("out must be a CUDA tensor on the same device")
        if tuple(out.shape) != (prepared.n_rows,) or out.dtype != torch.float32:
            raise ValueError("out shape/dtype must match result")
        if not out.is_contiguous():
            raise ValueError("out must be contiguous")
    if float(beta) != 0.0 and out is None:
        raise ValueError("beta != 0 requires out")
    y = out if out is not None else torch.empty(prepared.n_rows, dtype=torch.float32,
                                                 device=x.device)
    t0 = None
    if return_time:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
    sch = schedule or prepared.schedule
    if prepared.n_rows and sch.x_window != 0 and 0 < prepared.kernel_indices.numel() < 2**31:
        xw = _xwin_split(prepared, sch)
        _lib.xwin_gather(xw.layout, x, xw.xg, x.device)
        _lib.spmv_split(prepared._csr, FORMATS[prepared.value_format], prepared.hi, prepared.lo,
                        prepared.row_scale, xw.xg, y, alpha, beta,
                        xw.plan(_lib.FS_ACC_NATIVE).struct(1), x.device)
    elif prepared.n_rows and sch.x_slice_cols > 0 and prepared.n_cols > sch.x_slice_cols:
        sl = _sliced_split(prepared, sch)
        if sl.n_compact:
            _lib.spmv_split(sl.csr, FORMATS[prepared.value_format], sl.planes[0], sl.planes[1],
                            sl.row_scale, x, sl.z, 1.0, 0.0, sl.plan.struct(1), x.device)
        _lib.sliced_combine(sl.desc, sl.z, y, alpha, beta, x.device)
    elif prepared.n_rows:
        plan = prepared.plan(sch)
        _lib.spmv_split(prepared._csr, FORMATS[prepared.value_format], prepared.hi, prepared.lo,
                        prepared.row_scale, x, y, alpha, beta, plan.struct(1), x.device)
    if return_time:
        torch.cuda.synchronize()
        return y, (time.perf_counter() - t0) * 1000.0
    return y


__all__ = ["FORMATS", "PreparedCsrSpmvSplit", "prepare_spmv_csr_split",
           "flagsparse_spmv_csr_split", "decoded_values", "Schedule"]
