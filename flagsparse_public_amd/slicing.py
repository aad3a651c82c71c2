"""Column-sliced SpMV: each XCD's L2 holds only its own x columns.

A random x gather that misses L2 costs a full cache-line fill from the
Infinity Cache; with x far larger than one XCD's 4 MB L2, those fills, not
HBM, bound the SpMV.  Here the columns are cut into P = 8q slices of at most
``slice_cols`` columns and the matrix is stored slice-major: the entries of
slice k come before those of slice k + 1, grouped by row inside a slice.
Every (slice, row) pair with entries is a *compact row*.

One SpMV over the compact rows (entry tiles) writes a partial z[i] per compact
row.  Its tiles are dealt to 8 block groups -- group g takes the tiles of
slices [g*q, (g+1)*q) in order -- and blocks are dispatched to the XCDs
round-robin, so XCD g gathers only from its own x columns, one slice at a
time.  ``fs_sliced_combine`` then adds a row's partials in slice order (fp64)
and applies alpha / beta.  The result is deterministic.

Prepare cost: one stable sort of the entries by (slice, row) on the device,
and a second copy of col / values in slice-major order (the prepared object
keeps both layouts).
"""

from __future__ import annotations

import math

import torch

from . import _lib
from .rowsplit import RowSplitPlan, Schedule
from .sortkey import stable_order

N_GROUPS = 8          # MI355X: 8 XCDs, workgroups dealt round-robin
COMBINE_ROWS = 256    # rows per block of the combine pass
SLICED_DTYPES = (torch.float32, torch.float64, torch.complex64, torch.complex128)


def slice_count(n_cols: int, slice_cols: int) -> int:
    """P = 8q slices of at most slice_cols columns (q >= 1)."""
    q = max(1, math.ceil(n_cols / (N_GROUPS * max(int(slice_cols), 1))))
    return N_GROUPS * q


def row_masks(ukeys: torch.Tensor, n: int, n_slices: int, n_blocks: int) -> torch.Tensor:
    """Presence bits of the combine: word (k * n_blocks + b) * 8 + j, bit i
    set when row 256 b + 32 j + i has a compact row in slice k.  ``ukeys`` =
    k * n + row of every compact row (distinct), as int32 words."""
    words = COMBINE_ROWS // 32
    acc = torch.zeros(n_slices * n_blocks * words, dtype=torch.int64, device=ukeys.device)
    if ukeys.numel():
        k = ukeys // max(n, 1)
        r = ukeys % max(n, 1)
        idx = (k * n_blocks + r // COMBINE_ROWS) * words + (r % COMBINE_ROWS) // 32
        acc.index_add_(0, idx, torch.bitwise_left_shift(torch.ones_like(r), r % 32))
    # distinct bits per word, so the sum is the OR; store as two's-complement int32
    acc = torch.where(acc >= 2**31, acc - 2**32, acc)
    return acc.to(torch.int32)


class SlicedCsr:
    """Slice-major copy of one CSR matrix plus its compact-row plan."""

    def __init__(self, data, indices, indptr, n_rows, n_cols, slice_cols, schedule,
                 accumulate, planes=(), value_dtype=None, sums=False, fp32_sums=False):
        """``data`` may be None when the values live in other per-entry
        ``planes`` (split-precision storage, value_dtype float32); every plane
        is permuted into slice-major order with the entries.  ``sums``: the
        partials are the kernels' raw fp64 row sums (complex: two fp64)
        instead of values rounded to the value dtype -- the default fp32 path,
        which accumulates in fp64 and rounds once.  ``fp32_sums``: the combine
        adds fp32 partials in compensated fp32 instead of fp64 (the
        split-precision modes, which use no fp64)."""
        value_dtype = data.dtype if data is not None else value_dtype
        if value_dtype not in SLICED_DTYPES:
            raise TypeError(f"column slicing supports {SLICED_DTYPES}, got {value_dtype}")
        dev = indices.device
        self.n_rows, self.n_cols = int(n_rows), int(n_cols)
        n = self.n_rows
        P = slice_count(self.n_cols, slice_cols)
        self.n_slices = P
        self.width = max(1, math.ceil(self.n_cols / P))
        w = self.width
        nnz = indices.numel()
        lens = (indptr[1:] - indptr[:-1]).to(torch.int64)
        rows = torch.repeat_interleave(torch.arange(n, device=dev, dtype=torch.int64), lens)
        sl = indices.to(torch.int64) // w
        key = sl * n + rows
        del rows
        order = stable_order(key, P * n)
        key = key[order]
        self.cols = indices[order].contiguous()
        self.vals = data[order].contiguous() if data is not None else None
        self.planes = [pl[order].contiguous() if pl is not None else None for pl in planes]
        del order
        ukeys, counts = torch.unique_consecutive(key, return_counts=True)
        del key
        self.n_compact = int(ukeys.numel())
        ptr64 = torch.zeros(self.n_compact + 1, dtype=torch.int64, device=dev)
        torch.cumsum(counts, 0, out=ptr64[1:])
        self.cptr = ptr64.to(torch.int32) if nnz <= 2**31 - 1 else ptr64
        self.compact_row = (ukeys % max(n, 1)).to(torch.int32)
        # first entry of each slice (entries are slice-major): the entry offset
        # of the slice's first compact row
        first_compact = torch.searchsorted(
            ukeys, torch.arange(P + 1, device=dev, dtype=torch.int64) * max(n, 1), side="left")
        slice_entry = ptr64[first_compact]
        # combine bounds: first compact row of slice k in row block b
        n_blocks = -(-n // COMBINE_ROWS)
        q = (torch.arange(P, device=dev, dtype=torch.int64)[:, None] * n
             + torch.arange(n_blocks + 1, device=dev, dtype=torch.int64)[None, :] * COMBINE_ROWS)
        q = torch.minimum(q, (torch.arange(P, device=dev, dtype=torch.int64)[:, None] + 1) * n)
        self.bounds = torch.searchsorted(ukeys, q.reshape(-1).contiguous(), side="left")
        self.row_mask = row_masks(ukeys, n, P, n_blocks)
        del ukeys
        # entry tiles over the compact rows, dealt to XCD groups by slice
        T = schedule.tile_nnz if (schedule.tile_nnz > 0 and schedule.tile_mode == 1) else 2048
        self.schedule = Schedule(tile_nnz=T, tile_mode=1)
        self.plan = RowSplitPlan(self.cptr, self.n_compact, value_dtype, self.schedule, accumulate)
        qn = P // N_GROUPS
        # group g starts at the tile holding the first entry of slice g*q
        starts = slice_entry[torch.arange(0, P, qn, device=dev)].contiguous()
        first = torch.searchsorted(self.plan.tile_entry[1:].contiguous(), starts, side="right")
        gt = torch.empty(N_GROUPS + 1, dtype=torch.int64, device=dev)
        gt[:N_GROUPS] = torch.clamp(first, max=self.plan.n_tiles)
        gt[N_GROUPS] = self.plan.n_tiles
        gt[0] = 0
        self.group_tiles = gt
        self.group_blocks = int((gt[1:] - gt[:-1]).max().item())
        self.plan.set_groups(N_GROUPS, self.group_blocks, self.group_tiles)
        self.csr = _lib.FsCsr(value_type=_lib.value_code(value_dtype),
                              index_type=_lib.index_code(self.cols.dtype),
                              ptr_type=_lib.index_code(self.cptr.dtype), reserved=0,
                              n_rows=self.n_compact, n_cols=self.n_cols, nnz=nnz,
                              row_ptr=_lib.ptr(self.cptr), col_idx=_lib.ptr(self.cols),
                              values=_lib.ptr(self.vals))
        self.sums = bool(sums)
        z_dtype = value_dtype
        if self.sums:
            z_dtype = torch.complex128 if value_dtype.is_complex else torch.float64
        self.z = torch.empty(max(self.n_compact, 1), dtype=z_dtype, device=dev)
        if self.sums:
            self.plan.set_row_sums(self.z)
        self.desc = _lib.FsSlices(value_type=_lib.value_code(value_dtype), n_slices=P,
                                  n_rows=n, n_compact=self.n_compact,
                                  compact_row=_lib.ptr(self.compact_row),
                                  bounds=_lib.ptr(self.bounds), rows_per_block=COMBINE_ROWS,
                                  z_type=_lib.value_code(z_dtype),
                                  row_mask=_lib.ptr(self.row_mask),
                                  fp32_sums=int(bool(fp32_sums)), reserved=0)

    def run(self, x, y, alpha, beta, device):
        """y = alpha A x + beta y: compact-row partials, then the combine."""
        if self.n_compact > 0:
            _lib.spmv(self.csr, x, None if self.sums else self.z, 1.0, 0.0, self.plan.struct(1),
                      device)
        _lib.sliced_combine(self.desc, self.z, y, alpha, beta, device)
