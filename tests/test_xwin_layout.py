"""The x-window layout, checked on the CPU by replaying both passes in numpy.

Tile-major xg, pass 1 as the gather kernel runs it: entry k (window-major)
finds its run from its batch's first run plus the popcount of the batch's
run-start bits below it, and writes x[window * W + wcol[k]] to
xg[k + run_off[run]].  Window-major xg: pass 1 writes entry k to xg[k], and
pass 2 finds tile position d's run the same way and reads xg[d + run_off[run]].
Pass 2 as the entry kernel reads it: CSR entry j of tile t takes tile position
tile_entry[t] + x_perm[j].  With distinct x values, every entry must read
exactly x[col[j]], the destinations (sources) must be a permutation, and each
gather part must start inside its recorded window.  Then sum(val * xg) per
row equals the oracle's SpMV.
"""

import numpy as np
import pytest
import torch

from flagsparse_public_amd import synthetic
from flagsparse_public_amd.xwin import XWinCsr, part_ranges


def _part_cover(pw, pe, ws, nnz):
    """The gather kernels' walk over parts (csrc/xwin.hip): part p starts in
    window pw[p] at entry pe[p] and runs through windows until pe[p + 1].
    Returns, per entry, how many parts visited it and the window it was
    staged from, and the number of window stagings."""
    seen = np.zeros(nnz, np.int64)
    staged_win = np.full(nnz, -1, np.int64)
    stagings = 0
    n_windows = len(ws) - 1
    for p in range(len(pw)):
        k0, kend, w = int(pe[p]), int(pe[p + 1]), int(pw[p])
        while k0 < kend and w < n_windows:
            k1 = min(int(ws[w + 1]), kend)
            if k1 > k0:
                stagings += 1
                seen[k0:k1] += 1
                staged_win[k0:k1] = w
                k0 = k1
            w += 1
    return seen, staged_win, stagings


def _replay(xw, x):
    nnz = xw.nnz
    wcol = xw.wcol.numpy().view(np.uint16).astype(np.int64)
    br = xw.batch_run.numpy().astype(np.int64)
    bm = xw.batch_mask.numpy().view(np.uint64)
    off = xw.run_off.numpy().astype(np.int64)
    pw = xw.part_window.numpy().astype(np.int64)
    pe = xw.part_entry.numpy()
    ws = xw.window_entry.numpy()
    assert pe[0] == 0 and pe[-1] == nnz and np.all(np.diff(pe) > 0)
    assert ws[0] == 0 and ws[-1] == nnz and np.all(np.diff(ws) >= 0)
    assert np.all((ws[pw] <= pe[:-1]) & (pe[:-1] < ws[pw + 1]))  # first window of each part
    seen, staged_win, _ = _part_cover(pw, pe, ws, nnz)
    k = np.arange(nnz, dtype=np.int64)
    assert np.all(seen == 1)  # every entry gathered by exactly one part
    assert np.array_equal(staged_win, np.searchsorted(ws, k, side="right") - 1)
    lane = (k % 64).astype(np.uint64)
    below = (np.uint64(2) << lane) - np.uint64(2)  # lane 63: 2 << 63 wraps to 0
    r = br[k // 64] + np.bitwise_count(bm[k // 64] & below).astype(np.int64)
    assert r.min() >= 0 and r.max() < xw.n_runs
    moved = k + off[r]  # tile-major: destination of k; window-major: source of d = k
    assert np.array_equal(np.sort(moved), k), "run offsets do not give a permutation"
    win = np.searchsorted(ws, k, side="right") - 1  # the window of each entry
    gathered = x[win * xw.window + wcol]
    if xw.order == "tile":
        xg = np.empty(nnz, dtype=x.dtype)
        xg[moved] = gathered
        return xg  # already in tile order
    # pass 2 stages tile t from its own runs tile_run[t] .. tile_run[t+1]-1,
    # at most T of them
    te = xw.tile_entry.numpy()
    tr = xw.tile_run.numpy().astype(np.int64)
    assert tr[0] == 0 and tr[-1] == xw.n_runs and np.all(np.diff(tr) >= 0)
    assert np.all(np.diff(tr) <= xw.tile_nnz)
    tile = np.repeat(np.arange(xw.n_tiles), np.diff(te))
    assert np.all((tr[tile] <= r) & (r < tr[tile + 1]))
    return gathered[moved]  # the tile-order values pass 2 stages


def _check(ip, cols, vals, T, W, orc, order="window"):
    n_rows = ip.size - 1
    n_cols = int(cols.max()) + 1 if cols.size else 1
    n_cols = max(n_cols, 7 * W // 2 + 3)  # a last window narrower than W
    xw = XWinCsr(torch.from_numpy(vals), torch.from_numpy(cols), torch.from_numpy(ip),
                 n_rows, n_cols, tile_nnz=T, window=W, order=order)
    assert (xw.xg_runs is not None) == (order == "window")
    x = np.arange(n_cols, dtype=np.float64) + 0.25  # distinct
    xg = _replay(xw, x) if xw.nnz else np.zeros(0)
    te = xw.tile_entry.numpy()
    perm = xw.x_perm.numpy().view(np.uint16).astype(np.int64)
    tile = np.repeat(np.arange(xw.n_tiles), np.diff(te))
    assert perm.size == xw.nnz or xw.nnz == 0
    if xw.nnz:
        assert perm.max() < T
        got = xg[te[tile] + perm]
        assert np.array_equal(got, x[cols.astype(np.int64)])
        # each tile's x_perm is a permutation of its slots
        for t in range(0, xw.n_tiles, max(1, xw.n_tiles // 7)):
            p = np.sort(perm[te[t]:te[t + 1]])
            assert np.array_equal(p, np.arange(te[t + 1] - te[t]))
    # the two passes give the oracle's SpMV
    lens = np.diff(ip)
    rows = np.repeat(np.arange(n_rows), lens)
    y = np.zeros(n_rows)
    if xw.nnz:
        np.add.at(y, rows, vals.astype(np.float64) * xg[te[tile] + perm])
    y64 = orc.spmv(ip, cols, vals.astype(np.float64), x)
    assert np.allclose(y, y64, rtol=1e-12, atol=1e-9)
    return xw


@pytest.mark.parametrize("order", ["tile", "window"])
@pytest.mark.parametrize("T", [1024, 2048])
@pytest.mark.parametrize("W", [64, 1000, 4096])
def test_powerlaw_layout(orc, T, W, order):
    ip, cols, vals = synthetic.powerlaw_csr(6000, 5, alpha=1.8, max_len=5000)
    _check(ip, cols, vals, T, W, orc, order)


def test_row_capped_tiles_start_off_batch(orc):
    # runs of one-entry rows cut tiles every 768 rows at entry offsets that
    # are not multiples of 64: pass 2's positions then straddle batches
    lens = np.ones(9000, np.int64)
    lens[::97] = 40
    ip = np.zeros(lens.size + 1, np.int64)
    np.cumsum(lens, out=ip[1:])
    rng = np.random.default_rng(8)
    cols = rng.integers(0, 20_000, int(ip[-1])).astype(np.int32)
    vals = rng.standard_normal(cols.size).astype(np.float32)
    for order in ("tile", "window"):
        xw = _check(ip, cols, vals, 2048, 1000, orc, order)
        assert np.any(xw.tile_entry.numpy()[1:-1] % 64 != 0)


def test_banded_and_empty_rows(orc):
    ip, cols, vals = synthetic.banded_csr(5000, 3, 9)
    lens = np.diff(ip)
    lens[::3] = 0  # every third row empty
    keep = np.repeat(np.arange(lens.size), np.diff(ip)) % 3 != 0
    ip2 = np.zeros_like(ip)
    np.cumsum(lens, out=ip2[1:])
    for order in ("tile", "window"):
        _check(ip2, cols[keep], vals[keep], 2048, 256, orc, order)


def test_int64_indices_and_single_window(orc):
    ip, cols, vals = synthetic.poisson_csr(3000, 9.0, 4, index_dtype=np.int64)
    for order in ("tile", "window"):
        _check(ip.astype(np.int64), cols, vals, 1024, 4096, orc, order)
        _check(ip.astype(np.int64), cols, vals, 1024, 4, orc, order)  # windows of four columns


def test_empty_matrix():
    ip = np.zeros(11, np.int32)
    xw = XWinCsr(torch.zeros(0), torch.zeros(0, dtype=torch.int32), torch.from_numpy(ip), 10,
                 100, window=64)
    assert xw.nnz == 0 and xw.n_parts == 0


def test_part_ranges_cover_windows():
    ws = [0, 10, 10, 50_000, 50_001, 200_000]
    pw, pe = part_ranges(ws, 200_000, 16)
    assert pe[0] == 0 and pe[-1] == 200_000 and len(pe) == 17
    assert all(c % 64 == 0 for c in pe[:-1])  # cuts on batch boundaries
    assert max(np.diff(pe)) - min(np.diff(pe)) <= 64  # equal parts
    for p, w in enumerate(pw):
        assert ws[w] <= pe[p] < ws[w + 1]  # the window holding the first entry
    assert 1 not in pw  # the empty window holds no entry
    pw, pe = part_ranges([0, 100], 100, 16)  # fewer batches than parts
    assert pe == [0, 64, 100] and pw == [0, 0]


def test_part_ranges_window_aligned():
    """Window-major gather parts with at least two parts per non-empty
    window: each inside one window, dealt in proportion to the windows'
    entries, so each part stages one window --
    fewer stagings than equal ranges, which stage one more window per
    boundary they cross; every entry still gathered once from its window."""
    rng = np.random.default_rng(3)
    counts = rng.integers(30_000, 50_000, size=23)
    counts[[4, 11]] = 0  # empty windows get no part
    counts[7] = 1_000  # a small window gets one part
    ws = np.concatenate([[0], np.cumsum(counts)]).tolist()
    nnz = ws[-1]
    n_parts = 64
    pw, pe = part_ranges(ws, nnz, n_parts, window_aligned=True)
    pw, pe = np.asarray(pw), np.asarray(pe)
    assert pe[0] == 0 and pe[-1] == nnz and np.all(np.diff(pe) > 0)
    assert len(pw) == n_parts  # every share >= 1: exactly n_parts parts
    ws_a = np.asarray(ws)
    assert np.all((ws_a[pw] <= pe[:-1]) & (pe[1:] <= ws_a[pw + 1]))  # inside one window
    inner = pe[:-1][pe[:-1] != ws_a[pw]]
    assert np.all(inner % 64 == 0)  # cuts inside a window on 64-entry batches
    for w in range(len(counts)):
        sizes = np.diff(pe)[pw == w]
        if counts[w] == 0:
            assert sizes.size == 0
            continue
        # equal shares, each cut rounded up to a batch boundary
        assert sizes.sum() == counts[w] and sizes.max() - sizes.min() <= 128
        # parts in proportion to the window's entries (largest remainders)
        share = counts[w] * n_parts / nnz
        assert int(share) <= sizes.size <= max(1, int(share) + 1)
    seen, staged_win, stagings = _part_cover(pw, pe, ws, nnz)
    k = np.arange(nnz)
    assert np.all(seen == 1)
    assert np.array_equal(staged_win, np.searchsorted(ws_a, k, side="right") - 1)
    assert stagings == n_parts
    upw, upe = part_ranges(ws, nnz, n_parts)
    assert _part_cover(upw, upe, ws, nnz)[2] > stagings
    # fewer than two parts per non-empty window (21 of them): the equal ranges
    assert part_ranges(ws, nnz, 8, window_aligned=True) == part_ranges(ws, nnz, 8)
    assert part_ranges(ws, nnz, 41, window_aligned=True) == part_ranges(ws, nnz, 41)
    assert part_ranges(ws, nnz, 42, window_aligned=True) != part_ranges(ws, nnz, 42)
