"""Exhaustive schedule autotuning and its timing database.

The variant space is a fixed, ordered list of ``Schedule`` objects per
operation (SpMV, SpMM with 64 dense columns).  ``tune_spmv`` / ``tune_spmm``
time every variant on one prepared matrix with HIP events (median of
``reps`` timed launches after ``warmup``) and return the times; the
smallest is the exhaustive choice.  The learned selector is trained on
these timings and judged against them (its regret is time(selected) /
time(exhaustive best)).

``TimingDB`` keeps (features, per-variant times) rows and saves them as an
``.npz`` keyed by device name and variant-list version, so collection can
resume.
"""

from __future__ import annotations

import dataclasses
import os
import statistics

import numpy as np
import torch

from .rowsplit import Schedule

# Version of each operation's variant list: a timing database or a model
# recorded for another list is not used.
VARIANTS_VERSION = {"spmv": 6, "spmm64": 4}

# SpMV (fp32, native accumulation) variants.  0-5: entry tiles (fixed runs of
# entries), 3-5 with the columns cut into L2-sized slices when x is larger
# than one slice; 6-8: row tiles with chunked long rows; 9-15: one lane group
# per row (S lanes), long rows chunked; 16-17: entry tiles of 2048 / 4096
# over x gathered from LDS-staged windows of 128 KB (window-major xg).
SPMV_VARIANTS = (
    Schedule(tile_nnz=1024, tile_mode=1),
    Schedule(tile_nnz=2048, tile_mode=1),
    Schedule(tile_nnz=4096, tile_mode=1),
    Schedule(tile_nnz=2048, tile_mode=1, x_slice_cols=512 * 1024),
    Schedule(tile_nnz=4096, tile_mode=1, x_slice_cols=256 * 1024),
    Schedule(tile_nnz=2048, tile_mode=1, x_slice_cols=256 * 1024),
    Schedule(group_lanes=8, short_max=2048, chunk_nnz=8192, unroll=4, tile_nnz=2048),
    Schedule(group_lanes=8, short_max=1024, chunk_nnz=4096, unroll=8, tile_nnz=1024),
    Schedule(group_lanes=8, short_max=256, chunk_nnz=2048, unroll=4, tile_nnz=512),
    Schedule(group_lanes=1, short_max=64, chunk_nnz=512, unroll=4),
    Schedule(group_lanes=2, short_max=64, chunk_nnz=512, unroll=4),
    Schedule(group_lanes=4, short_max=64, chunk_nnz=512, unroll=8),
    Schedule(group_lanes=8, short_max=128, chunk_nnz=512, unroll=4),
    Schedule(group_lanes=16, short_max=256, chunk_nnz=512, unroll=4),
    Schedule(group_lanes=32, short_max=512, chunk_nnz=1024, unroll=4),
    Schedule(group_lanes=64, short_max=4096, chunk_nnz=4096, unroll=8),
    Schedule(tile_nnz=2048, tile_mode=1, x_window=-1),
    Schedule(tile_nnz=4096, tile_mode=1, x_window=-1),
)
# The plain default-tile SpMV the learned selector is compared with: a whole
# wavefront per row over fixed blocks of entries -- the structure of the
# reference's default path, one program per row stepping through 256-entry
# blocks (spmv_csr.py:158-184, 664-708).
SPMV_DEFAULT_INDEX = 15
# A tuned fixed schedule reported beside it: 8-lane groups per row.
SPMV_FIXED_S8_INDEX = 12


def spmv_heuristic_index(n_cols: int, n_rows: int | None = None, nnz: int | None = None,
                         item: int = 4, local: bool | None = None) -> int:
    """Variant chosen by the fixed rule of rowsplit.default_spmv_schedule for
    fp32 (``item`` 4) or fp64 (8) values.  x larger than 8 MB: with the
    matrix's column locality known (``local``, rowsplit.columns_local), plain
    entry tiles of 2048 (variant 1) for local columns and the x-window tiles
    of 4096 (17) otherwise; without it, entry tiles over 2 MB column slices
    (3 for fp32, 5 for fp64).  Else row groups of width pow2ceil(mean row
    length / 2) (the nearest variant)."""
    if n_cols * item > (8 << 20):
        if local is not None:
            return 1 if local else 17
        return 3 if item == 4 else 5
    if n_rows is None or nnz is None:
        return SPMV_FIXED_S8_INDEX
    from .rowsplit import rowgroup_spmv_schedule
    s = rowgroup_spmv_schedule(n_rows, nnz).group_lanes
    return {1: 9, 2: 10, 4: 11, 8: 12, 16: 13, 32: 14, 64: 15}[s]


# SpMM (64 dense columns) variants: column tile (lane-group width), short-row
# bound, chunk size, XCD remap.
SPMM_VARIANTS = (
    Schedule(short_max=256, chunk_nnz=256, col_tile=64),
    Schedule(short_max=256, chunk_nnz=256, col_tile=32),
    Schedule(short_max=256, chunk_nnz=256, col_tile=16),
    Schedule(short_max=64, chunk_nnz=128, col_tile=64),
    Schedule(short_max=1024, chunk_nnz=1024, col_tile=64),
    Schedule(short_max=256, chunk_nnz=256, col_tile=64, xcd_remap=1),
    Schedule(short_max=256, chunk_nnz=256, col_tile=32, xcd_remap=1),
    Schedule(short_max=64, chunk_nnz=128, col_tile=32, xcd_remap=1),
)
SPMM_DEFAULT_INDEX = 0


HOLD_US = 1500.0  # device time the queue is held while the host enqueues a group
_hold_cycles = None


def _hold():
    """Cycles of torch.cuda._sleep that keep the stream busy for about
    HOLD_US, measured once per process."""
    global _hold_cycles
    if _hold_cycles is None:
        probe = 2_000_000
        torch.cuda._sleep(probe)  # warm the spin kernel
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        stream = torch.cuda.current_stream()
        a.record(stream)
        torch.cuda._sleep(probe)
        b.record(stream)
        b.synchronize()
        us = max(a.elapsed_time(b) * 1000.0, 1.0)
        _hold_cycles = max(1, int(probe * HOLD_US / us))
    return _hold_cycles


def time_launches(fn, reps: int = 7, warmup: int = 2, inner: int = 1) -> float:
    """Median device microseconds per call of fn() over reps event-timed
    groups of ``inner`` calls on the current stream.  Each group is enqueued
    behind a short spin kernel, so the launches run back to back and the
    timing is the device's, not the host's launch rate."""
    for _ in range(warmup):
        fn()
    stream = torch.cuda.current_stream()
    hold = _hold()
    samples = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        torch.cuda._sleep(hold)
        a.record(stream)
        for _ in range(inner):
            fn()
        b.record(stream)
        b.synchronize()
        samples.append(a.elapsed_time(b) * 1000.0 / inner)
    return float(statistics.median(samples))


def tune_spmv(prepared, x, variants=SPMV_VARIANTS, reps=7, warmup=2, inner=1):
    """Median us of every SpMV variant (native accumulation) on one matrix;
    each variant runs as what is built for the matrix's dtypes
    (rowsplit.supported_spmv_schedule)."""
    from .rowsplit import supported_spmv_schedule
    from .sparse_operations.spmv_csr import _hip_spmv_csr_impl_opt_prepared
    y = torch.empty(prepared.n_rows, dtype=prepared.data.dtype, device=prepared.data.device)
    times = []
    dt, it = prepared.data.dtype, prepared.kernel_indices.dtype
    for sch in (supported_spmv_schedule(v, dt, it, prepared.data.numel()) for v in variants):
        times.append(time_launches(
            lambda s=sch: _hip_spmv_csr_impl_opt_prepared(prepared, x, out=y, schedule=s),
            reps, warmup, inner))
    return np.asarray(times)


def tune_spmm(prepared, B, variants=SPMM_VARIANTS, reps=5, warmup=1, inner=1):
    """Median us of every SpMM variant (native accumulation) on one matrix,
    each launched as the library launches a schedule it chose itself (entries
    in flight and column tile adjusted for B's dtype and width)."""
    from .sparse_operations.spmm_csr import (_hip_spmm_csr_impl_opt_prepared,
                                             _library_spmm_launch)
    C = torch.empty((prepared.n_rows, B.shape[1]), dtype=B.dtype, device=B.device)
    n_dense = int(B.shape[1])
    times = []
    for sch in variants:
        run, col_tile = _library_spmm_launch(sch, B.dtype, n_dense)
        run = dataclasses.replace(run, col_tile=col_tile)
        times.append(time_launches(
            lambda s=run: _hip_spmm_csr_impl_opt_prepared(prepared, B, out=C, schedule=s),
            reps, warmup, inner))
    return np.asarray(times)


def best_schedule(times, variants):
    i = int(np.argmin(times))
    return i, variants[i]


class TimingDB:
    """Rows of (matrix id, family, features[16], times[K]) for one operation."""

    def __init__(self, op: str, n_variants: int, device_name: str = ""):
        self.op = op
        self.n_variants = n_variants
        self.device_name = device_name
        self.ids, self.families, self.features, self.times, self.sizes = [], [], [], [], []

    def add(self, mid: int, family: str, feats, times, n_rows: int, nnz: int):
        self.ids.append(int(mid))
        self.families.append(family)
        self.features.append(np.asarray(feats, np.float64))
        self.times.append(np.asarray(times, np.float64))
        self.sizes.append((int(n_rows), int(nnz)))

    def __len__(self):
        return len(self.ids)

    def arrays(self):
        return (np.asarray(self.ids, np.int64), np.asarray(self.families),
                np.asarray(self.features).reshape(-1, 16),
                np.asarray(self.times).reshape(-1, self.n_variants),
                np.asarray(self.sizes, np.int64).reshape(-1, 2))

    def save(self, path: str):
        ids, fam, feats, times, sizes = self.arrays()
        np.savez_compressed(path, op=self.op, device_name=self.device_name,
                            version=VARIANTS_VERSION[self.op], ids=ids, families=fam,
                            features=feats, times=times, sizes=sizes)

    @classmethod
    def load(cls, path: str) -> "TimingDB":
        z = np.load(path, allow_pickle=False)
        db = cls(str(z["op"]), int(z["times"].shape[1]), str(z["device_name"]))
        want = VARIANTS_VERSION[db.op]
        if int(z["version"]) != want:
            raise ValueError(f"{path}: variant list version {int(z['version'])} != {want}")
        for i, f, ft, t, sz in zip(z["ids"], z["families"], z["features"], z["times"], z["sizes"]):
            db.add(int(i), str(f), ft, t, int(sz[0]), int(sz[1]))
        return db


def default_db_dir() -> str:
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "selector_models")
