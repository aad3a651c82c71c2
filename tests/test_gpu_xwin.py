"""GPU parity of the x-window SpMV (x gathered from LDS windows, then entry
tiles over the gathered values).

Both passes keep the entry-tile kernel's products and summation order, so the
x-window result must equal the one-pass entry-tile result with the same tile
size bit for bit, for every accumulation mode; and both must meet the
oracle's bars (fp32 default path: <= 1 ulp of fl32(y64) and >= 99.9 %
bit-identical rows; fp32 native: 1e-5 of sum|a x| per row; fp64: 1e-12).
"""

import numpy as np
import pytest
import torch

from flagsparse_public_amd import _lib, flagsparse_spmv_csr, prepare_spmv_csr, synthetic
from flagsparse_public_amd.rowsplit import RowSplitPlan, Schedule
from flagsparse_public_amd.sparse_operations.spmv_csr import XWIN_WIDE_TILE
from flagsparse_public_amd.xwin import XWinCsr
from tests._util import host64, row_rel_err, to_dev, ulp_distance_f32

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _matrix(kind, dtype):
    npd = np.float64 if dtype == torch.float64 else np.float32
    if kind == "powerlaw":
        ip, cols, vals = synthetic.powerlaw_csr(200_000, 21, alpha=1.9, max_len=60_000, dtype=npd)
    elif kind == "banded":
        ip, cols, vals = synthetic.banded_csr(150_000, 22, 17, dtype=npd)
    else:  # poisson rows with a run of empty rows and columns past the last window
        ip, cols, vals = synthetic.poisson_csr(90_000, 14.0, 23, dtype=npd, n_cols=170_003)
        lens = np.diff(ip)
        lens[20_000:26_000] = 0
        keep = np.ones(cols.size, bool)
        keep[ip[20_000]:ip[26_000]] = False
        ip = np.zeros_like(ip)
        np.cumsum(lens, out=ip[1:])
        cols, vals = cols[keep], vals[keep]
    n_cols = 170_003 if kind == "poisson" else ip.size - 1
    return ip, cols, vals, n_cols


def _oracle(orc, ip, cols, vals, x):
    return orc.spmv(ip, cols, vals, x), orc.spmv_abs(ip, cols, vals, x)


def _check(y, y64, cond, dtype, use_opt):
    y = host64(y)
    if dtype == torch.float32 and not use_opt:
        ulps = ulp_distance_f32(y.astype(np.float32), y64.astype(np.float32))
        assert ulps.max() <= 1 and (ulps == 0).mean() >= 0.999
    else:
        tol = 1e-5 if dtype == torch.float32 else 1e-12
        assert row_rel_err(y, y64, cond).max() <= tol


CASES = [(torch.float32, True, 1024), (torch.float32, True, 2048), (torch.float32, True, 4096),
         (torch.float32, False, 2048), (torch.float32, False, 4096),
         (torch.float64, True, 1024), (torch.float64, True, 2048)]


@pytest.mark.parametrize("dtype,use_opt,T", CASES)
@pytest.mark.parametrize("kind", ["powerlaw", "banded", "poisson"])
@pytest.mark.parametrize("window", [-1, 4096])
def test_xwin_equals_entry_tiles(orc, kind, dtype, use_opt, T, window):
    ip, cols, vals, n_cols = _matrix(kind, dtype)
    n = ip.size - 1
    print(f"[xwin] {kind} n={n} nnz={vals.size} {dtype} opt={use_opt} T={T} W={window}",
          flush=True)
    x = synthetic.dense_uniform(5, n_cols, dtype=vals.dtype)
    args = [to_dev(vals), to_dev(cols), to_dev(ip)]
    # the default fp32 path runs x-window layouts with tiles of at most
    # XWIN_WIDE_TILE entries: the entry tiles it equals are cut the same way
    T_ref = min(T, XWIN_WIDE_TILE) if dtype == torch.float32 and not use_opt else T
    ref = prepare_spmv_csr(*args, (n, n_cols), schedule=Schedule(tile_nnz=T_ref, tile_mode=1))
    xw = prepare_spmv_csr(*args, (n, n_cols),
                          schedule=Schedule(tile_nnz=T, tile_mode=1, x_window=window))
    dx = to_dev(x)
    y_ref = flagsparse_spmv_csr(x=dx, prepared=ref, use_opt=use_opt)
    y = flagsparse_spmv_csr(x=dx, prepared=xw, use_opt=use_opt)
    assert torch.equal(y, y_ref)
    y2 = flagsparse_spmv_csr(x=dx, prepared=xw, use_opt=use_opt)  # reuses the layout
    assert torch.equal(y2, y)
    layouts = [v for k, v in xw._plans.items() if isinstance(k, tuple) and k[0] == "xwin"]
    assert len(layouts) == 1 and layouts[0].tile_nnz == T_ref
    y64, cond = _oracle(orc, ip, cols, vals, x)
    _check(y, y64, cond, dtype, use_opt)


@pytest.mark.parametrize("dtype,T", [(torch.float32, 1024), (torch.float32, 2048),
                                     (torch.float32, 4096), (torch.float64, 2048)])
@pytest.mark.parametrize("kind", ["powerlaw", "poisson"])
@pytest.mark.parametrize("window", [128, 4096, None])
def test_xg_orders_equal_entry_tiles(kind, dtype, T, window):
    """Tile-major and window-major xg give the one-pass entry tiles' result
    bit for bit (windows of 128 columns give power-law tiles of 2048 or 4096
    entries more than 1024 runs, which the staging loads in a loop)."""
    ip, cols, vals, n_cols = _matrix(kind, dtype)
    n = ip.size - 1
    print(f"[xwin-order] {kind} {dtype} T={T} W={window}", flush=True)
    x = to_dev(synthetic.dense_uniform(8, n_cols, dtype=vals.dtype))
    args = [to_dev(vals), to_dev(cols), to_dev(ip)]
    ref = prepare_spmv_csr(*args, (n, n_cols), schedule=Schedule(tile_nnz=T, tile_mode=1))
    y_ref = flagsparse_spmv_csr(x=x, prepared=ref, use_opt=True)
    for order in ("tile", "window"):
        xw = XWinCsr(*args, n, n_cols, tile_nnz=T, window=window, order=order)
        if window == 128 and order == "window" and kind == "powerlaw" and T >= 2048:
            assert int((xw.tile_run[1:] - xw.tile_run[:-1]).max()) > 1024
        y = torch.empty_like(y_ref)
        xw.run(x, y, 1.0, 0.0, _lib.FS_ACC_NATIVE, x.device)
        assert torch.equal(y, y_ref), order


def test_xwin_alpha_beta_int64_and_misaligned_x(orc):
    ip, cols, vals, n_cols = _matrix("powerlaw", torch.float32)
    n = ip.size - 1
    x = synthetic.dense_uniform(6, n_cols + 1)
    dxx = to_dev(x)[1:]  # 4-byte offset: the window staging takes its scalar path
    y0 = synthetic.dense_uniform(7, n)
    for idx_dtype in (np.int32, np.int64):
        args = [to_dev(vals), to_dev(cols.astype(idx_dtype)), to_dev(ip.astype(idx_dtype))]
        p = prepare_spmv_csr(*args, (n, n_cols),
                             schedule=Schedule(tile_nnz=2048, tile_mode=1, x_window=-1))
        out = to_dev(y0)
        res = flagsparse_spmv_csr(x=dxx, prepared=p, use_opt=True, out=out, alpha=0.5, beta=-2.0)
        assert res is out
        y64, cond = _oracle(orc, ip, cols, vals, x[1:])
        want = 0.5 * y64 - 2.0 * y0.astype(np.float64)
        err = row_rel_err(host64(res), want, 0.5 * cond + 2.0 * np.abs(y0))
        assert err.max() <= 1e-5


@pytest.mark.slow
def test_xwin_powerlaw_4m_full_size(orc):
    """The 4M-row power-law matrix (48.6M entries, x of 16 MB): the x-window
    result equals the one-pass entry tiles and meets the oracle's bars."""
    ip, cols, vals = synthetic.spmv_powerlaw_4m()
    n = ip.size - 1
    print(f"[xwin] powerlaw_4m nnz={vals.size}", flush=True)
    x = synthetic.dense_uniform(2, n)
    args = [to_dev(vals), to_dev(cols), to_dev(ip)]
    y64, cond = _oracle(orc, ip, cols, vals, x)
    dx = to_dev(x)
    ref = prepare_spmv_csr(*args, (n, n), schedule=Schedule(tile_nnz=2048, tile_mode=1))
    xw = prepare_spmv_csr(*args, (n, n),
                          schedule=Schedule(tile_nnz=2048, tile_mode=1, x_window=-1))
    for use_opt in (True, False):
        y = flagsparse_spmv_csr(x=dx, prepared=xw, use_opt=use_opt)
        assert torch.equal(y, flagsparse_spmv_csr(x=dx, prepared=ref, use_opt=use_opt))
        _check(y, y64, cond, torch.float32, use_opt)


@pytest.mark.parametrize("case", range(64))
def test_xwin_random_shapes(case):
    """Seeded random shapes: sizes from a few rows to 60k, densities from
    mostly-empty to ~60 entries per row with runs of one-entry rows (row-cap
    tile cuts off 64-entry batches), windows from 4 to 65,532 columns, tile
    sizes 1024 / 2048 / 4096, both xg orders, fp32 / fp64, int32 / int64
    indices: bit-identical to the one-pass entry tiles."""
    rng = np.random.default_rng(1000 + case)
    dtype = (torch.float32, torch.float64)[case % 2]
    npd = np.float32 if dtype == torch.float32 else np.float64
    n = int(np.exp(rng.uniform(np.log(5), np.log(60_000))))
    n_cols = int(np.exp(rng.uniform(np.log(4), np.log(300_000))))
    lens = rng.poisson(rng.uniform(0.05, 60), n)
    if case % 3 == 0:
        lens[: n // 2] = 1  # one-entry rows: tiles cut every 768 rows
    lens = np.minimum(lens, n_cols)
    ip = np.zeros(n + 1, np.int64)
    np.cumsum(lens, out=ip[1:])
    nnz = int(ip[-1])
    cols = rng.integers(0, n_cols, nnz)
    vals = rng.uniform(-1, 1, nnz).astype(npd)
    idx_dtype = (np.int32, np.int64)[(case // 2) % 2]
    T = (1024, 2048, 4096)[case % 3] if dtype == torch.float32 else (1024, 2048)[case % 2]
    max_w = 32768 if dtype == torch.float32 else 16384
    W = int(rng.choice([4, 64, 1000, 4096, max_w]))
    W -= W % 4
    order = ("tile", "window")[(case // 4) % 2]
    print(f"[xwin-rand] case {case}: n={n} n_cols={n_cols} nnz={nnz} {dtype} T={T} W={W} "
          f"{order}", flush=True)
    args = [to_dev(vals), to_dev(cols.astype(idx_dtype)), to_dev(ip.astype(idx_dtype))]
    x = to_dev(rng.uniform(-1, 1, n_cols).astype(npd))
    # the one-pass entry tiles launched directly with native accumulation, on
    # int32 copies of the indices (an int64-index prepared matrix takes the
    # default path even with use_opt, as in the reference -- supports_opt needs
    # int32 columns -- and one-pass tiles of 1024 / 4096 are built for int32
    # columns only; the index type does not enter the arithmetic)
    a32 = [args[0], args[1].to(torch.int32), args[2].to(torch.int32)]
    plan = RowSplitPlan(a32[2], n, dtype, Schedule(tile_nnz=T, tile_mode=1), _lib.FS_ACC_NATIVE)
    y_ref = torch.empty(n, dtype=dtype, device=x.device)
    _lib.spmv(_lib.make_csr(*a32, n, n_cols), x, y_ref, 1.0, 0.0, plan.struct(1), x.device)
    if nnz == 0:
        return
    xw = XWinCsr(*args, n, n_cols, tile_nnz=T, window=W, order=order)
    y = torch.empty_like(y_ref)
    xw.run(x, y, 1.0, 0.0, _lib.FS_ACC_NATIVE, x.device)
    assert torch.equal(y, y_ref)
