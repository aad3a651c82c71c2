"""Column-sliced SpMV: keep the x gathers inside L2.

A matrix whose x does not fit an XCD's L2 (4 MB on MI355X) pays one L2 miss
-- a full cache-line fill from the Infinity Cache -- for most of its 4-byte
x gathers.  Cutting the columns into P slices of at most ``slice_cols``
columns and running one launch per slice keeps every launch's x slice
resident in every XCD's L2: launch k computes y = alpha A_k x_k + beta' y with
beta' = beta for k = 0 and 1 afterwards.  The cost is P row-pointer reads and
P - 1 extra read-modify-writes of y, paid in streaming bytes instead of
fabric round trips.

The per-slice matrices are built once (stable sort of the entries by
(slice, row), column indices made slice-relative), so entries of a row keep
their order inside each slice and the result is deterministic.  Used for the
native-accumulation path (fp32 partial sums per slice, one fp32 add per
slice).
"""

from __future__ import annotations

import math

import torch

from . import _lib
from .rowsplit import RowSplitPlan


class SlicedCsr:
    """Per-slice CSR matrices and plans of one matrix (see module doc)."""

    def __init__(self, data, indices, indptr, n_rows, n_cols, slice_cols, schedule,
                 accumulate):
        dev = data.device
        self.n_rows, self.n_cols = int(n_rows), int(n_cols)
        self.n_slices = max(1, math.ceil(self.n_cols / max(int(slice_cols), 1)))
        self.width = math.ceil(self.n_cols / self.n_slices) if self.n_cols else 0
        w, P, n = self.width, self.n_slices, self.n_rows
        nnz = data.numel()
        lens = (indptr[1:] - indptr[:-1]).to(torch.int64)
        rows = torch.repeat_interleave(torch.arange(n, device=dev, dtype=torch.int64), lens)
        cols = indices.to(torch.int64)
        sl = cols // w if w else torch.zeros_like(cols)
        key = sl * n + rows
        order = torch.argsort(key, stable=True)
        del rows
        sl_sorted = sl[order]
        self.cols = (cols[order] - sl_sorted * w).to(torch.int32)
        self.vals = data[order].contiguous()
        del cols, sl, order, sl_sorted
        counts = torch.bincount(key, minlength=P * n).view(P, n)
        del key
        ptr_dtype = torch.int32 if nnz <= 2**31 - 1 else torch.int64
        self.parts = []
        start = 0
        for k in range(P):
            ip = torch.zeros(n + 1, dtype=torch.int64, device=dev)
            torch.cumsum(counts[k], 0, out=ip[1:])
            nk = int(ip[-1].item())
            ip = ip.to(ptr_dtype)
            c0 = k * w
            wk = max(0, min(w, self.n_cols - c0))
            cols_k = self.cols[start:start + nk]
            vals_k = self.vals[start:start + nk]
            csr = _lib.make_csr(vals_k, cols_k, ip, n, wk)
            plan = RowSplitPlan(ip, n, data.dtype, schedule, accumulate,
                                row_lengths=counts[k].to(torch.int64))
            self.parts.append((csr, plan, ip, c0, wk))
            start += nk

    def run(self, x, y, alpha, beta, device):
        """y = alpha A x + beta y, one launch (plus fix-up) per slice."""
        for k, (csr, plan, _ip, c0, wk) in enumerate(self.parts):
            xk = x[c0:c0 + wk]
            _lib.spmv(csr, xk, y, alpha, beta if k == 0 else 1.0, plan.struct(1), device)
