"""Row-split schedules and their device metadata (the ``fs_plan`` of the C ABI).

A schedule sends every row of length <= ``short_max`` to the short-row role
(groups of ``group_lanes`` lanes, natural row order) and cuts every longer row
into chunks of ``chunk_nnz`` entries reduced by one wavefront each.  The plan
holds, in device memory: the long-row ids, each long row's first chunk, the
long-row ordinal of every chunk and the chunk-partial workspace (read by the
second, fix-up launch for rows that span several chunks).

This replaces the reference's row-length buckets (spmv_csr.py:128-153,
345-391; spmm_csr.py:395-477; spmm_csr_opt_alg2.py:60-166, 307-335): the
short role needs no row list at all, and the long-row split is computed on
the device instead of the host loop of spmm_csr.py:417-441.
"""

from __future__ import annotations

import dataclasses

import torch

from . import _lib


@dataclasses.dataclass(frozen=True)
class Schedule:
    group_lanes: int = 8      # SpMV lanes per short row (1..64, power of two)
    short_max: int = 128      # longest row of the short role
    chunk_nnz: int = 512      # entries per long-row chunk (multiple of 64)
    unroll: int = 4           # entries in flight per lane
    xcd_remap: int = 0        # 1: contiguous row ranges per XCD
    col_tile: int = 0         # SpMM dense columns per lane group (0: all)
    tile_nnz: int = 0         # SpMV: > 0 selects the tile role for short rows
    x_slice_cols: int = 0     # SpMV native path: > 0 cuts the columns into slices of
                              # at most this many (one launch per slice, x slice in L2)
    tile_mode: int = 0        # SpMV with tile_nnz > 0: 0 row tiles, 1 entry tiles
    x_window: int = 0         # SpMV fp32 / fp64 with entry tiles: != 0 gathers x from
                              # windows staged in LDS first (columns per window; -1: 128 KB)
    b_panels: int = 1         # SpMM: launches over column panels of B (1: all columns)

    def key(self) -> "Schedule":
        """Cache key (a frozen dataclass hashes and compares by its fields)."""
        return self


def _pow2_ceil(v: float) -> int:
    p = 1
    while p < v and p < 64:
        p <<= 1
    return p


TILE_ROWS = 512  # rows per tile at most (the kernel's LDS row-pointer slots)


L2_SLICE_BYTES = 2 << 20    # x columns per XCD slice: half of one XCD's 4 MB L2
SLICE_MIN_X_BYTES = 8 << 20  # slice only when x is well beyond what L2 + MALL serve well


LOCALITY_SAMPLE_ROWS = 4096  # rows sampled by far_column_fraction
LOCAL_X_BYTES = 128 << 10  # a row is local when its first column is this close to its diagonal


def far_column_fraction(indptr: torch.Tensor, indices: torch.Tensor, n_rows: int, n_cols: int,
                        far_cols: int) -> float | None:
    """Fraction of up to LOCALITY_SAMPLE_ROWS evenly spaced nonempty rows
    whose first column lies more than far_cols columns from the row's
    diagonal position (row * n_cols // n_rows); None when no sampled row has
    entries.  One small device pass and one host read."""
    if n_rows <= 0 or indices.numel() == 0:
        return None
    dev = indptr.device
    step = max(1, n_rows // LOCALITY_SAMPLE_ROWS)
    rows = torch.arange(0, n_rows, step, device=dev, dtype=torch.int64)
    start = indptr[rows].to(torch.int64)
    nonempty = indptr[rows + 1].to(torch.int64) > start
    first = indices[torch.clamp(start, max=indices.numel() - 1)].to(torch.int64)
    diag = rows * n_cols // n_rows
    far = ((first - diag).abs() > far_cols) & nonempty
    counts = torch.stack([far.sum(), nonempty.sum()]).cpu().tolist()
    return counts[0] / counts[1] if counts[1] else None


def columns_local(indptr: torch.Tensor, indices: torch.Tensor, n_rows: int, n_cols: int,
                  item: int) -> bool | None:
    """Whether most sampled rows start within LOCAL_X_BYTES of x (values of
    ``item`` bytes) of their diagonal: banded, block-diagonal and
    bandwidth-reduced matrices are; random column patterns (uniform,
    power-law, R-MAT) are not.  None for a matrix without entries."""
    far = far_column_fraction(indptr, indices, n_rows, n_cols, max(1, LOCAL_X_BYTES // item))
    return None if far is None else far < 0.5


def default_spmv_schedule(n_rows: int, nnz: int, value_dtype=None, n_cols: int | None = None,
                          indptr: torch.Tensor | None = None,
                          indices: torch.Tensor | None = None) -> Schedule:
    """Fixed rule used when no selector applies.  x larger than 8 MB: when
    the matrix's row pointers and columns are given, a sample of rows
    decides (columns_local) between plain entry tiles of 2048 (local
    columns: the x gathers hit L2) and, for random columns, the x-window
    layout with tiles of 4096 (fp32 / fp64 values, up to XWIN_MAX_NNZ
    entries) or entry tiles of 2048 over the column-sliced layout (2 MB of x
    per slice; other values, larger matrices, or without the matrix's
    arrays).  Otherwise one lane group per row, its
    width from the mean row length (rowgroup_spmv_schedule)."""
    if value_dtype is not None and n_cols is not None:
        item = torch.empty(0, dtype=value_dtype).element_size()
        if n_cols * item > SLICE_MIN_X_BYTES:
            sliced = Schedule(tile_nnz=2048, tile_mode=1, x_slice_cols=L2_SLICE_BYTES // item)
            if indptr is None or indices is None:
                return sliced
            if columns_local(indptr, indices, n_rows, n_cols, item):
                return Schedule(tile_nnz=2048, tile_mode=1)
            if value_dtype in (torch.float32, torch.float64) and nnz <= XWIN_MAX_NNZ:
                return Schedule(tile_nnz=4096, tile_mode=1, x_window=-1)
            return sliced
    return rowgroup_spmv_schedule(n_rows, nnz)


XWIN_MAX_NNZ = 2**31 - 1  # entries the x-window layout addresses


def supported_spmv_schedule(sch: Schedule, value_dtype, index_dtype,
                            nnz: int | None = None) -> Schedule:
    """The schedule a library-chosen variant runs as for these dtypes: entry
    tiles of 1024 / 4096 entries are built for fp32 / fp64 values with int32
    columns only (others: tiles of 2048), and the x-window layout (whose second
    pass indexes through 16-bit positions, any column type) for fp32 / fp64
    values and at most XWIN_MAX_NNZ entries (``nnz``, when given) -- beyond
    that the launch runs plain entry tiles of the same size."""
    if sch.tile_mode != 1 or sch.tile_nnz in (0, 2048):
        return sch
    fp = value_dtype in (torch.float32, torch.float64)
    xwin = sch.x_window != 0 and (nnz is None or nnz <= XWIN_MAX_NNZ)
    if fp and (xwin or index_dtype == torch.int32):
        return sch
    return dataclasses.replace(sch, tile_nnz=2048)


def row_tile_spmv_schedule(n_rows: int, nnz: int, value_dtype=None) -> Schedule:
    """Row tiles of ~2048 entries for rows up to 2048 entries (1024 for
    complex values, whose fp64 row sums take 16 bytes of LDS each), longer
    rows in 8192-entry chunks."""
    t = 1024 if (value_dtype is not None and value_dtype.is_complex) else 2048
    return Schedule(group_lanes=8, short_max=t, chunk_nnz=8192, unroll=4, tile_nnz=t)


def rowgroup_spmv_schedule(n_rows: int, nnz: int) -> Schedule:
    """The lane-group-per-row schedule: group width near half the mean row
    length, rows over 16 entries per lane chunked in 512-entry chunks."""
    mean = nnz / n_rows if n_rows > 0 else 0.0
    s = _pow2_ceil(max(mean / 2.0, 1.0))
    return Schedule(group_lanes=s, short_max=max(16 * s, 64), chunk_nnz=512, unroll=4)


def tile_starts(row_lengths: torch.Tensor, short_max: int, tile_nnz: int,
                max_rows: int = TILE_ROWS) -> torch.Tensor:
    """First row of every tile, plus n_rows at the end (int64, on the device).

    Cuts: where the running count of short-row entries crosses a multiple of
    tile_nnz, every max_rows rows, and around every row longer than
    short_max (which becomes a one-row tile that the tile role skips).  A
    tile then holds < tile_nnz + short_max entries and <= max_rows rows.
    """
    dev = row_lengths.device
    n = row_lengths.numel()
    if n == 0:
        return torch.zeros(1, dtype=torch.int64, device=dev)
    lens = row_lengths.to(torch.int64)
    long_mask = lens > short_max
    eff = torch.where(long_mask, torch.zeros_like(lens), lens)
    cum_excl = torch.cumsum(eff, 0) - eff
    key = cum_excl // int(tile_nnz)
    parts = [torch.zeros(1, dtype=torch.int64, device=dev),
             torch.nonzero(key[1:] != key[:-1], as_tuple=False).flatten() + 1,
             torch.arange(0, n, max_rows, dtype=torch.int64, device=dev),
             torch.full((1,), n, dtype=torch.int64, device=dev)]
    longs = torch.nonzero(long_mask, as_tuple=False).flatten()
    if longs.numel():
        parts += [longs, longs + 1]
    return torch.unique(torch.cat(parts))


def default_spmm_schedule(n_rows: int, nnz: int, n_dense: int) -> Schedule:
    return Schedule(group_lanes=1, short_max=256, chunk_nnz=256, unroll=4, col_tile=0)


class RowSplitPlan:
    """Device metadata of one schedule for one matrix (see module docstring).

    Calls that share a plan must be ordered on one stream: the chunk
    partials are reused by every call.
    """

    def __init__(self, indptr: torch.Tensor, n_rows: int, value_dtype, schedule: Schedule,
                 accumulate: int, row_lengths: torch.Tensor | None = None):
        device = indptr.device
        self.schedule = schedule
        self.accumulate = int(accumulate)
        self.value_code = _lib.value_code(value_dtype)
        self.device = device
        self._work = {}
        self._structs = {}
        self.n_groups, self.group_blocks, self.group_tiles = 0, 0, None
        self.row_sums = None
        self.x_perm = None
        self.xg_runs = None
        if schedule.tile_nnz > 0 and schedule.tile_mode == _lib.FS_TILE_ENTRIES:
            self._init_entry_tiles(indptr, n_rows, schedule)
            return
        if row_lengths is None:
            row_lengths = (indptr[1:].to(torch.int64) - indptr[:-1].to(torch.int64)) if n_rows > 0 \
                else torch.zeros(0, dtype=torch.int64, device=device)
        long_mask = row_lengths > schedule.short_max
        self.long_rows = torch.nonzero(long_mask, as_tuple=False).flatten().to(torch.int64)
        self.n_long = int(self.long_rows.numel())
        if self.n_long:
            lens = row_lengths[self.long_rows]
            per = (lens + schedule.chunk_nnz - 1) // schedule.chunk_nnz
            self.long_chunk_ptr = torch.zeros(self.n_long + 1, dtype=torch.int64, device=device)
            torch.cumsum(per, 0, out=self.long_chunk_ptr[1:])
            self.n_chunks = int(self.long_chunk_ptr[-1].item())
            self.chunk_long = torch.repeat_interleave(
                torch.arange(self.n_long, dtype=torch.int32, device=device), per)
        else:
            self.long_chunk_ptr = torch.zeros(1, dtype=torch.int64, device=device)
            self.n_chunks = 0
            self.chunk_long = torch.zeros(0, dtype=torch.int32, device=device)
        if schedule.tile_nnz > 0:
            if schedule.short_max > schedule.tile_nnz:
                raise ValueError("tile role needs short_max <= tile_nnz")
            self.tile_start = tile_starts(row_lengths, schedule.short_max, schedule.tile_nnz)
            self.n_tiles = int(self.tile_start.numel()) - 1
        else:
            self.tile_start = None
            self.n_tiles = 0

    ENTRY_ROW_CAP = 768  # owned rows per entry tile (the kernel's LDS row batch)

    def _init_entry_tiles(self, indptr, n_rows, schedule):
        """Entry tiles: cut the entries every T and, inside a tile that owns
        more than ENTRY_ROW_CAP rows (runs of empty or one-entry rows), cut
        again every ENTRY_ROW_CAP rows at a row start.  Tile t holds entries
        [tile_entry[t], tile_entry[t+1]) and owns rows [tile_start[t],
        tile_start[t+1]), which start in that range (the last tile also owns
        the rows that start at nnz)."""
        device = indptr.device
        T = int(schedule.tile_nnz)
        cap = self.ENTRY_ROW_CAP
        ptr = indptr.to(torch.int64)
        nnz = int(ptr[-1].item()) if n_rows > 0 else 0
        nb = max(1, -(-nnz // T))
        R = torch.zeros(nb + 1, dtype=torch.int64, device=device)
        R[-1] = n_rows
        if nb > 1:
            bounds = torch.arange(1, nb, dtype=torch.int64, device=device) * T
            R[1:-1] = torch.searchsorted(ptr[:-1].contiguous(), bounds, side="left")
        E = torch.arange(nb + 1, dtype=torch.int64, device=device) * T
        E[-1] = nnz
        counts = R[1:] - R[:-1]
        extra = torch.clamp((counts + cap - 1) // cap - 1, min=0)
        n_extra = int(extra.sum().item())
        if n_extra:
            base = torch.repeat_interleave(torch.arange(nb, device=device), extra)
            first = torch.cumsum(extra, 0) - extra          # first extra of each base tile
            k = torch.arange(n_extra, device=device) - first[base] + 1
            R_x = R[base] + k * cap
            E_x = ptr[R_x]
            # boundaries are monotone in both coordinates: order by (entry, row)
            key = torch.cat([E[:-1], E_x]) * (n_rows + 1) + torch.cat([R[:-1], R_x])
            order = torch.argsort(key, stable=True)
            R = torch.cat([torch.cat([R[:-1], R_x])[order], R[-1:]])
            E = torch.cat([torch.cat([E[:-1], E_x])[order], E[-1:]])
        self.n_tiles = int(R.numel()) - 1
        self.tile_start = R.contiguous()
        self.tile_entry = E.contiguous()
        self.long_rows = torch.zeros(0, dtype=torch.int64, device=device)
        self.long_chunk_ptr = torch.zeros(1, dtype=torch.int64, device=device)
        self.chunk_long = torch.zeros(0, dtype=torch.int32, device=device)
        self.n_long = 0
        self.n_chunks = 0
        self.entry_tiles = True

    def set_groups(self, n_groups: int, group_blocks: int, group_tiles: torch.Tensor):
        """Entry tiles: deal tile ranges to block groups (block b serves group
        b % n_groups); group_tiles[g] .. group_tiles[g + 1] must cover every
        tile once and group_blocks be the largest range."""
        self.n_groups = int(n_groups)
        self.group_blocks = int(group_blocks)
        self.group_tiles = group_tiles
        self._structs.clear()

    def set_row_sums(self, sums: torch.Tensor):
        """Entry tiles: store raw fp64 row sums here instead of writing y."""
        self.row_sums = sums
        self._structs.clear()

    def set_x_perm(self, perm: torch.Tensor, runs=None):
        """Entry tiles: x is gathered per entry (x windows); entry j of tile t
        reads tile position d = tile_entry[t] + perm[j], at x[d] (tile-major
        x) or, with ``runs`` = (tile_run, batch_run, batch_mask, run_off) of
        the tile positions, at x[d + run_off[run of d]] (window-major x)."""
        self.x_perm = perm
        self.xg_runs = runs
        self._structs.clear()

    def struct(self, n_dense: int = 1) -> _lib.FsPlan:
        """fs_plan for a call with n_dense dense columns (1 for SpMV)."""
        key = max(int(n_dense), 1)
        st = self._structs.get(key)
        if st is not None:
            return st
        if getattr(self, "entry_tiles", False):
            if key != 1:
                raise ValueError("entry tiles are an SpMV schedule")
            nbytes = _lib.tile_carry_bytes(self.value_code, self.n_tiles)
            carries = torch.empty(max(nbytes, 8), dtype=torch.uint8, device=self.device)
            self._work[key] = carries
            sch = self.schedule
            st = _lib.FsPlan(accumulate=self.accumulate, group_lanes=sch.group_lanes,
                             xcd_remap=0, unroll=sch.unroll, short_max=0, chunk_nnz=64,
                             n_long=0, long_rows=_lib.ptr(self.long_rows),
                             long_chunk_ptr=_lib.ptr(self.long_chunk_ptr), n_chunks=0,
                             chunk_long=_lib.ptr(self.chunk_long), partials=carries.data_ptr(),
                             partials_bytes=int(carries.numel()), tile_nnz=sch.tile_nnz,
                             n_tiles=self.n_tiles, tile_start=_lib.ptr(self.tile_start),
                             tile_mode=_lib.FS_TILE_ENTRIES, n_groups=self.n_groups,
                             group_blocks=self.group_blocks,
                             group_tiles=_lib.ptr(self.group_tiles),
                             row_sums=_lib.ptr(self.row_sums),
                             tile_entry=_lib.ptr(self.tile_entry),
                             x_perm=_lib.ptr(self.x_perm))
            runs = self.xg_runs
            if runs is not None:
                st.xg_tile_run = _lib.ptr(runs[0])
                st.xg_batch_run = _lib.ptr(runs[1])
                st.xg_batch_mask = _lib.ptr(runs[2])
                st.xg_run_off = _lib.ptr(runs[3])
                st.xg_n_runs = int(runs[3].numel())
            self._structs[key] = st
            return st
        if self.n_long:
            nbytes = _lib.partials_bytes(self.value_code, self.accumulate, self.n_chunks, key)
            partials = torch.empty(max(nbytes, 8), dtype=torch.uint8, device=self.device)
        else:
            partials = torch.empty(8, dtype=torch.uint8, device=self.device)
        self._work[key] = partials
        sch = self.schedule
        st = _lib.FsPlan(
            accumulate=self.accumulate,
            group_lanes=sch.group_lanes,
            xcd_remap=sch.xcd_remap,
            unroll=sch.unroll,
            short_max=sch.short_max,
            chunk_nnz=sch.chunk_nnz,
            n_long=self.n_long,
            long_rows=_lib.ptr(self.long_rows),
            long_chunk_ptr=_lib.ptr(self.long_chunk_ptr),
            n_chunks=self.n_chunks,
            chunk_long=_lib.ptr(self.chunk_long),
            partials=partials.data_ptr(),
            partials_bytes=int(partials.numel()),
            tile_nnz=sch.tile_nnz if key == 1 else 0,
            n_tiles=self.n_tiles if key == 1 else 0,
            tile_start=_lib.ptr(self.tile_start) if (key == 1 and self.tile_start is not None) else None,
            tile_mode=_lib.FS_TILE_ROWS,
        )
        self._structs[key] = st
        return st
