"""fs_spmv_csr_noalloc through the C ABI: the same results as fs_spmv_csr
with a caller-owned partials buffer, bit for bit (the same kernels in the same
order), for plans whose kernels keep partials (entry-tile carries, long-row
chunk partials) and one that keeps none; two calls in a row, the second on a
smaller matrix, so the second runs over the scratch the first left behind.
(The refusal of a plan beyond the scratch is checked on the CPU in
test_abi.py: it returns before anything touches the device.)"""

import ctypes

import numpy as np
import pytest
import torch

from flagsparse_public_amd import _lib, synthetic
from flagsparse_public_amd.rowsplit import RowSplitPlan, Schedule
from tests._util import host64, row_rel_err, to_dev

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)

SCHEDULES = {
    "entry_tiles_2048": Schedule(tile_nnz=2048, tile_mode=_lib.FS_TILE_ENTRIES),
    "entry_tiles_1024": Schedule(tile_nnz=1024, tile_mode=_lib.FS_TILE_ENTRIES),
    "row_groups_chunked": Schedule(group_lanes=8, short_max=64, chunk_nnz=512, unroll=4),
    "row_tiles_chunked": Schedule(group_lanes=8, short_max=2048, chunk_nnz=8192, unroll=4,
                                  tile_nnz=2048),
    "row_groups_no_long": Schedule(group_lanes=16, short_max=1 << 40, chunk_nnz=512),
}


def _problem(n, seed, dtype):
    ip, cols, vals = synthetic.powerlaw_csr(n, seed, alpha=2.0, max_len=min(n, 20_000),
                                            dtype=np.float64 if dtype == torch.float64 else np.float32)
    x = synthetic.dense_uniform(seed + 1, n, dtype=vals.dtype)
    y0 = synthetic.dense_uniform(seed + 2, n, dtype=vals.dtype)
    return ip, cols, vals, x, y0


def _noalloc(csr, x, y, alpha, beta, plan):
    lib = _lib.load()
    with _lib.on_device(DEV):
        return lib.fs_spmv_csr_noalloc(ctypes.byref(csr), _lib.ptr(x), _lib.ptr(y), float(alpha),
                                       float(beta), ctypes.byref(plan), _lib.stream_handle(DEV))


@pytest.mark.parametrize("dtype,accumulate", [(torch.float32, _lib.FS_ACC_NATIVE),
                                              (torch.float32, _lib.FS_ACC_WIDE),
                                              (torch.float64, _lib.FS_ACC_NATIVE)])
@pytest.mark.parametrize("name", sorted(SCHEDULES))
def test_noalloc_equals_fs_spmv_csr(orc, name, dtype, accumulate):
    sch = SCHEDULES[name]
    for n, seed in ((120_000, 11), (9_000, 12)):  # second call smaller
        ip, cols, vals, x, y0 = _problem(n, seed, dtype)
        print(f"[noalloc] {name} {dtype} acc={accumulate} n={n} nnz={vals.size}", flush=True)
        d_ip, d_cols, d_vals = to_dev(ip), to_dev(cols), to_dev(vals)
        d_x = to_dev(x)
        csr = _lib.make_csr(d_vals, d_cols, d_ip, n, n)
        plan = RowSplitPlan(d_ip, n, dtype, sch, accumulate)
        st = plan.struct()
        need = (_lib.tile_carry_bytes(csr.value_type, plan.n_tiles)
                if sch.tile_mode == _lib.FS_TILE_ENTRIES and sch.tile_nnz else
                _lib.partials_bytes(csr.value_type, accumulate, plan.n_chunks))
        assert need <= _lib.FS_NOALLOC_SCRATCH_BYTES
        for alpha, beta in ((1.0, 0.0), (0.75, -1.5)):
            ref = to_dev(y0)
            _lib.spmv(csr, d_x, ref, alpha, beta, st, DEV)
            got = to_dev(y0)
            rc = _noalloc(csr, d_x, got, alpha, beta, st)
            assert rc == _lib.FS_OK, _lib.load().fs_last_error()
            torch.cuda.synchronize()
            assert torch.equal(got, ref), (name, n, alpha, beta)
        # and against the oracle (alpha = 0.75, beta = -1.5 from the last call)
        y64 = orc.spmv(ip, cols, vals, x)
        cond = orc.spmv_abs(ip, cols, vals, x)
        want = 0.75 * y64 - 1.5 * y0.astype(np.float64)
        tol = 1e-5 if dtype == torch.float32 else 1e-12
        err = row_rel_err(host64(got), want, 0.75 * cond + 1.5 * np.abs(y0))
        assert err.max() <= tol



@pytest.mark.parametrize("order", ["tile", "window"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_noalloc_x_window_plans(order, dtype):
    """The gathered-x form (x-window second pass, both xg orders): the carries
    go to the scratch array, the results equal fs_spmv_csr's bit for bit."""
    from flagsparse_public_amd.xwin import XWinCsr
    for n, seed in ((150_000, 21), (20_000, 22)):  # second call smaller
        ip, cols, vals, x, y0 = _problem(n, seed, dtype)
        print(f"[noalloc] x-window {order} {dtype} n={n} nnz={vals.size}", flush=True)
        d_ip, d_cols, d_vals, d_x = to_dev(ip), to_dev(cols), to_dev(vals), to_dev(x)
        xw = XWinCsr(d_vals, d_cols, d_ip, n, n, tile_nnz=2048, window=4096, order=order)
        _lib.xwin_gather(xw.layout, d_x, xw.xg, DEV)
        st = xw.plan(_lib.FS_ACC_NATIVE).struct(1)
        ref = to_dev(y0)
        _lib.spmv(xw.csr, xw.xg, ref, 0.5, 2.0, st, DEV)
        got = to_dev(y0)
        rc = _noalloc(xw.csr, xw.xg, got, 0.5, 2.0, st)
        assert rc == _lib.FS_OK, _lib.load().fs_last_error()
        torch.cuda.synchronize()
        assert torch.equal(got, ref), (order, n)
