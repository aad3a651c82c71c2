"""x-window SpMV: random x gathers served from LDS instead of L2.

At ~12 entries per row over a 16 MB x, a one-pass SpMV issues one L2 request
per 4-byte x gather, and the L2 request rate, not HBM bandwidth, bounds it.
This schedule splits the SpMV into two streaming passes:

1. ``fs_xwin_gather``: the columns are cut into windows of W columns (one
   window of x fits one workgroup's LDS: W = 32768 fp32 / 16384 fp64 on
   MI355X's 160 KB).  A second copy of the column indices, stored
   window-major as 16-bit in-window columns, is streamed window by window;
   each entry's x value is read from LDS and written to ``xg`` in the
   second pass's layout.
2. ``fs_spmv_csr`` with entry tiles and ``x_perm``: tile t's gathered values
   lie at ``xg[tile_entry[t] .. tile_entry[t+1])``, ordered by window (the
   order pass 1 writes them in); the tile stages them in LDS and reads entry
   j's value through ``x_perm[j]``.  Values are read in CSR order, so the
   tile's products and row sums are exactly those of the one-pass entry-tile
   kernel with the same tiles: the results are bit-identical.

Layout (all on the device, built once per matrix by stable sorts):

* ``wcol[k]`` (uint16): in-window column of window-major entry k.  Window-
  major order is (window, tile, CSR order).
* runs: the entries of one tile inside one window.  ``run_off[r]``: entry k of
  run r goes to ``xg[k + run_off[r]]``.  ``batch_run[b]`` / ``batch_mask[b]``:
  run of entry 64 b and the entries of batch b that start a run, so the
  gather finds an entry's run with one popcount.
* ``x_perm[j]`` (uint16): position of CSR entry j inside its tile's slice of
  ``xg``.
* parts: the gather's workgroups, about one per CU.  For the window-major
  xg with at least two parts per non-empty window, each part lies inside
  one window (the parts dealt to the windows in proportion to their
  entries), so each part stages one window; otherwise they are equal
  ranges of the window-major entries cut on 64-entry batches, and a range
  runs through the windows it covers, staging each in turn.

Bytes per entry (fp32): pass 1 reads 2 (wcol) and writes 4 (xg); pass 2 reads
4 (value) + 4 (xg) + 2 (x_perm).  16 bytes, all in streams, against 12 in the
one-pass kernel of which 4 are random gathers.

``order`` places xg in one of two orders:

* ``"tile"`` (tile-major): xg is in the second pass's order, so pass 2 reads
  each tile's values as one stream, and pass 1 writes each run to its place
  -- runs of ~T / n_windows entries, neighbouring ones written by different
  workgroups, so most cache lines of xg are written in parts.
* ``"window"`` (window-major): pass 1 writes xg[k] for window-major entry k,
  whole lines in order, and the run table (``batch_run`` / ``batch_mask`` /
  ``run_off``, now over the tile-order positions, with ``tile_run`` = the
  first run of each tile) moves to pass 2, whose loads find tile position d
  at xg[d + run_off[run of d]].
"""

from __future__ import annotations

import math

import torch

from . import _lib
from .rowsplit import RowSplitPlan, Schedule
from .sortkey import stable_order, stable_sort

XWIN_DTYPES = (torch.float32, torch.float64)
DEFAULT_WINDOW_BYTES = 128 * 1024
ORDERS = {"tile": _lib.FS_XG_TILE_MAJOR, "window": _lib.FS_XG_WINDOW_MAJOR}
N_XCDS = 8  # MI355X: workgroups are dealt to 8 XCDs round-robin


def default_window(value_dtype, device) -> int:
    """Columns per window: 128 KB of x, within what one workgroup can stage."""
    item = torch.empty(0, dtype=value_dtype).element_size()
    w = DEFAULT_WINDOW_BYTES // item
    cap = _lib.xwin_max_window(_lib.value_code(value_dtype), device)
    if cap > 0:
        w = min(w, cap)
    return max(4, w - w % 4)


def _as_int16(t: torch.Tensor) -> torch.Tensor:
    """Integer values in [0, 65536) stored as their 16-bit pattern (the
    integer narrowing conversion keeps the low 16 bits)."""
    return t.to(torch.int16)


def run_table(is_start: torch.Tensor, off_at_start: torch.Tensor):
    """(batch_run, batch_mask, run_off) of a sequence cut into runs:
    ``is_start[i]`` marks the first position of each run (position 0
    included), ``off_at_start`` holds each run's offset, in run order.
    batch_run[b] = run of position 64 b; bit i >= 1 of batch_mask[b] is set
    when position 64 b + i starts a run."""
    dev, n = is_start.device, is_start.numel()
    nb = -(-n // 64)
    pad = torch.zeros(nb * 64, dtype=torch.uint8, device=dev)
    pad[:n] = is_start
    rows = pad.view(nb, 64)
    # run of position 64 b: the run starts before batch b, plus one if 64 b
    # starts a run, minus one (runs count from 0)
    starts = rows.sum(dim=1)
    batch_run = (torch.cumsum(starts, 0) - starts + rows[:, 0] - 1).to(torch.int32).contiguous()
    # bit 8 k + j of batch b from byte j of group k: pack 8 flags into a byte,
    # then 8 bytes into the word (distinct bits: each sum is the OR)
    b8 = torch.bitwise_left_shift(rows.view(nb, 8, 8).to(torch.int32),
                                  torch.arange(8, device=dev, dtype=torch.int32)).sum(dim=2)
    batch_mask = torch.bitwise_left_shift(
        b8, torch.arange(0, 64, 8, device=dev, dtype=torch.int64)).sum(dim=1).contiguous()
    return batch_run, batch_mask, off_at_start.to(torch.int32).contiguous()


def gather_parts(device, window_bytes: int) -> int:
    """Workgroups of the gather: one per CU for each window that fits the
    CU's LDS at once, so the whole gather runs in one round."""
    cus, lds = 256, 160 * 1024
    if isinstance(device, torch.device) and device.type == "cuda":
        props = torch.cuda.get_device_properties(device)
        cus = int(getattr(props, "multi_processor_count", cus))
    return cus * max(1, lds // max(window_bytes, 1))


def part_ranges(win_start, nnz: int, n_parts: int, window_aligned: bool = False):
    """Cut the window-major entries [0, nnz) into about n_parts ranges.
    Returns host lists (part_window, part_entry): the window holding each
    range's first entry, and the range starts plus nnz at the end.

    Default: n_parts about equal ranges cut on multiples of 64 (the
    tile-major gather's batch size); a range may run through several
    windows and stages each.  ``window_aligned`` (the window-major gather,
    which takes a range starting anywhere): when there are at least two
    parts per non-empty window, the parts are dealt to the windows in
    proportion to their entries (largest remainders first) and cut inside
    each window on multiples of 64 -- one window staged per part, n_parts
    stagings in all instead of n_parts plus one per window boundary that
    falls inside a range.  On the 4M-row power-law SpMV that is 2.3 us
    faster in fp32 (2.1 parts per window) and 8 us slower in fp64 (1.04
    parts per window: whole-window parts), hence the two-part bound."""
    import bisect
    ws = [int(v) for v in win_start]
    if window_aligned and nnz > 0:
        counts = [ws[w + 1] - ws[w] for w in range(len(ws) - 1)]
        full = [w for w, c in enumerate(counts) if c > 0]
        if 2 * len(full) <= n_parts:
            share = [counts[w] * n_parts / nnz for w in full]
            alloc = [max(1, int(s)) for s in share]
            spare = n_parts - sum(alloc)
            if spare > 0:
                by_rest = sorted(range(len(full)), key=lambda i: int(share[i]) - share[i])
                for i in by_rest[:spare]:
                    alloc[i] += 1
            pw, cuts = [], []
            for w, a in zip(full, alloc):
                lo, hi = ws[w], ws[w + 1]
                pw.append(w)
                cuts.append(lo)
                for j in range(1, a):
                    c = -(-(lo + (hi - lo) * j // a) // 64) * 64
                    if cuts[-1] < c < hi:
                        pw.append(w)
                        cuts.append(c)
            return pw, cuts + [nnz]
    cuts = [0]
    for i in range(1, max(int(n_parts), 1)):
        c = -(-(nnz * i // n_parts) // 64) * 64
        if cuts[-1] < c < nnz:
            cuts.append(c)
    pw = [bisect.bisect_right(ws, c) - 1 for c in cuts]
    return pw, cuts + [nnz]


class XWinCsr:
    """The x-window layout of one CSR matrix, its gathered-x buffer and the
    second pass's entry-tile plans (one per accumulation mode)."""

    def __init__(self, data, indices, indptr, n_rows, n_cols, tile_nnz=2048, window=None,
                 value_dtype=None, order="window"):
        """``data`` may be None when the values are stored elsewhere (split-
        precision planes): the layout depends only on the columns and row
        pointers, and x / xg have ``value_dtype``.  ``order``: "window" or
        "tile" (see the module doc)."""
        if order not in ORDERS:
            raise ValueError(f"order must be one of {sorted(ORDERS)}, got {order!r}")
        self.order = order
        value_dtype = data.dtype if data is not None else value_dtype
        if value_dtype not in XWIN_DTYPES:
            raise TypeError(f"x windows support {XWIN_DTYPES}, got {value_dtype}")
        dev = indices.device
        nnz = int(indices.numel())
        if nnz > 2**31 - 1:
            raise ValueError("x windows address at most 2^31 - 1 entries")
        self.data, self.indices, self.indptr = data, indices, indptr
        self.n_rows, self.n_cols, self.nnz = int(n_rows), int(n_cols), nnz
        self.device = dev
        self.value_dtype = value_dtype
        self.tile_nnz = int(tile_nnz)
        W = int(window) if window else default_window(value_dtype, dev)
        if W <= 0 or W % 4 or W > 65532:
            raise ValueError(f"window must be a multiple of 4 in (0, 65532], got {W}")
        self.window = W
        self.n_windows = max(1, -(-self.n_cols // W))
        self.schedule = Schedule(tile_nnz=self.tile_nnz, tile_mode=_lib.FS_TILE_ENTRIES)
        self._plans = {}
        base = RowSplitPlan(indptr, self.n_rows, value_dtype, self.schedule, _lib.FS_ACC_NATIVE)
        self._plans[_lib.FS_ACC_NATIVE] = base
        self.tile_entry = base.tile_entry
        self.n_tiles = base.n_tiles
        self._build(indices, base)
        self.xg = torch.empty(max(nnz, 1), dtype=value_dtype, device=dev)
        self.csr = (_lib.make_csr(data, indices, indptr, self.n_rows, self.n_cols)
                    if data is not None else None)

    def _build(self, indices, base):
        dev, nnz, W, S = self.device, self.nnz, self.window, self.n_windows
        n_tiles = self.n_tiles
        te = self.tile_entry
        self.xg_runs = None
        if nnz == 0:
            z16 = torch.zeros(1, dtype=torch.int16, device=dev)
            self.wcol, self.x_perm = z16, z16
            self.batch_run = torch.zeros(1, dtype=torch.int32, device=dev)
            self.batch_mask = torch.zeros(1, dtype=torch.int64, device=dev)
            self.run_off = torch.zeros(1, dtype=torch.int32, device=dev)
            self.n_runs = 0
            self.part_window = torch.zeros(1, dtype=torch.int32, device=dev)
            self.part_entry = torch.zeros(2, dtype=torch.int64, device=dev)
            self.window_entry = torch.zeros(S + 1, dtype=torch.int64, device=dev)
            self.n_parts = 0
            self.layout = self._struct()
            return
        # index arithmetic in int32 wherever the values fit (entry positions
        # do: nnz < 2^31), so each pass over the entries moves half the bytes
        it = torch.int32 if self.n_cols <= 2**31 - 1 else torch.int64
        cols = indices.to(it)
        win = cols // W
        tile = torch.repeat_interleave(torch.arange(n_tiles, device=dev, dtype=torch.int32),
                                       te[1:] - te[:-1])
        # window-major order (window, tile, CSR order): entries are already in
        # tile order, so a stable sort by window is enough
        order1 = stable_order(win, S)
        self.wcol = _as_int16((cols - win * W)[order1]).contiguous()
        del cols
        # pass-2 layout: by (tile, window, CSR order)
        kt = torch.int32 if n_tiles * S <= 2**31 - 1 else torch.int64
        key2, order2 = stable_sort(tile.to(kt) * S + win.to(kt), n_tiles * S)
        pos2 = torch.empty(nnz, dtype=torch.int32, device=dev)
        pos2[order2] = torch.arange(nnz, device=dev, dtype=torch.int32)
        self.x_perm = _as_int16(pos2 - te.to(torch.int32)[tile]).contiguous()
        # runs: the (window, tile) groups, contiguous in both orders
        if self.order == "tile":
            # over the window-major entries k: entry k goes to xg[k + run_off]
            del order2, key2
            key1 = (win.to(kt) * n_tiles + tile.to(kt))[order1]
            del tile
            is_start = torch.ones(nnz, dtype=torch.bool, device=dev)
            is_start[1:] = key1[1:] != key1[:-1]
            del key1
            run_k0 = torch.nonzero(is_start, as_tuple=False).flatten()
            dst = pos2[order1]
            del pos2
            self.batch_run, self.batch_mask, self.run_off = run_table(
                is_start, dst[run_k0] - run_k0)
            del dst, is_start
        else:
            # over the tile-order positions d: position d reads xg[d + run_off]
            del tile, pos2
            is_start = torch.ones(nnz, dtype=torch.bool, device=dev)
            is_start[1:] = key2[1:] != key2[:-1]
            run_d0 = torch.nonzero(is_start, as_tuple=False).flatten()
            # each run's first position in xg: window-major order lays the
            # runs out by (window, tile), so a run starts after the runs that
            # sort before it -- an exclusive sum of run lengths in that order
            rkey = key2[run_d0].to(torch.int64)
            del key2, order2
            lens = torch.diff(run_d0, append=torch.full((1,), nnz, dtype=torch.int64, device=dev))
            o = torch.argsort((rkey % S) * n_tiles + rkey // S)
            src = torch.empty_like(run_d0)
            src[o] = torch.cumsum(lens[o], 0) - lens[o]
            del rkey, lens, o
            self.batch_run, self.batch_mask, self.run_off = run_table(is_start, src - run_d0)
            del src, is_start
            # tile t's runs: those starting in [te[t], te[t+1]) (runs never
            # cross a tile boundary)
            self.tile_run = torch.searchsorted(run_d0, te).to(torch.int32).contiguous()
            del run_d0
            self.xg_runs = (self.tile_run, self.batch_run, self.batch_mask, self.run_off)
        self.n_runs = int(self.run_off.numel())
        # window ranges in window-major order (entries per window, summed up)
        # -> gather parts
        ws = torch.zeros(S + 1, dtype=torch.int64, device=dev)
        torch.cumsum(torch.bincount(win, minlength=S), 0, out=ws[1:])
        del win, order1
        self.window_entry = ws.contiguous()
        item = torch.empty(0, dtype=self.value_dtype).element_size()
        pw, pe = part_ranges(ws.cpu().tolist(), nnz, gather_parts(dev, W * item),
                             window_aligned=self.order == "window")
        self.part_window = torch.tensor(pw, dtype=torch.int32, device=dev)
        self.part_entry = torch.tensor(pe, dtype=torch.int64, device=dev)
        self.n_parts = len(pw)
        self.layout = self._struct()

    def _struct(self) -> _lib.FsXwin:
        return _lib.FsXwin(value_type=_lib.value_code(self.value_dtype), window=self.window,
                           n_cols=self.n_cols, nnz=self.nnz, n_runs=self.n_runs,
                           n_parts=self.n_parts, wcol=_lib.ptr(self.wcol),
                           batch_run=_lib.ptr(self.batch_run),
                           batch_mask=_lib.ptr(self.batch_mask), run_off=_lib.ptr(self.run_off),
                           part_window=_lib.ptr(self.part_window),
                           part_entry=_lib.ptr(self.part_entry),
                           window_entry=_lib.ptr(self.window_entry),
                           order=ORDERS[self.order], reserved=0)

    def plan(self, accumulate: int) -> RowSplitPlan:
        p = self._plans.get(int(accumulate))
        if p is None:
            p = RowSplitPlan(self.indptr, self.n_rows, self.value_dtype, self.schedule,
                             accumulate)
            self._plans[int(accumulate)] = p
        if p.x_perm is None:
            p.set_x_perm(self.x_perm, self.xg_runs)
            if self.xg_runs is not None and self.n_tiles >= 2 * N_XCDS:
                # neighbouring tiles read the two ends of shared cache lines
                # of xg (a tile's run in a window ends where the next tile's
                # begins): deal each XCD a contiguous tile range, so both
                # reads of a line go through one L2
                gt = torch.tensor([self.n_tiles * g // N_XCDS for g in range(N_XCDS + 1)],
                                  dtype=torch.int64, device=self.device)
                p.set_groups(N_XCDS, -(-self.n_tiles // N_XCDS), gt)
        return p

    def run(self, x, y, alpha, beta, accumulate, device):
        """y = alpha A x + beta y (x dense, n_cols values)."""
        if self.n_rows == 0:
            return
        if self.nnz > 0:
            _lib.xwin_gather(self.layout, x, self.xg, device)
        _lib.spmv(self.csr, self.xg, y, alpha, beta, self.plan(accumulate).struct(1), device)

    def footprint_bytes(self) -> int:
        """Device bytes this layout adds to the matrix."""
        t = (self.wcol, self.x_perm, self.batch_run, self.batch_mask, self.run_off,
             self.part_window, self.part_entry, self.window_entry, self.xg)
        if self.xg_runs is not None:
            t = t + (self.tile_run,)
        return int(sum(a.numel() * a.element_size() for a in t))


def window_count(n_cols: int, window: int) -> int:
    return max(1, math.ceil(n_cols / max(window, 1)))
