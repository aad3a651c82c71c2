"""GPU tests of the split-precision value storage modes (MI355X).

* Encoding: the bf16 hi / lo planes equal the oracle's round-to-nearest-even
  split bit for bit; the fp16 row-scale plane decodes to within 2^-11 of the
  row scale.
* Kernel parity: y equals the fp64 oracle applied to the decoded values within
  the fp32-accumulation bar 1e-5 * sum|a x| per row.
* Stated error against the original fp32 matrix (per row, / sum|a x|):
  bf16x2 <= 2^-16 + 1e-6, bf16 <= 2^-8 + 1e-6, f16_rowscale <= 2^-11 + 2^-14.
"""

import numpy as np
import pytest
import torch

from tests._util import host64, row_rel_err, to_dev

pytestmark = pytest.mark.gpu

from flagsparse_public_amd import synthetic  # noqa: E402
from flagsparse_public_amd.precision import (  # noqa: E402
    decoded_values,
    flagsparse_spmv_csr_split,
    prepare_spmv_csr_split,
)
from flagsparse_public_amd.rowsplit import Schedule  # noqa: E402

BOUND_VS_FP32 = {"bf16x2": 2.0 ** -16 + 1e-6, "bf16": 2.0 ** -8 + 1e-6,
                 "f16_rowscale": 2.0 ** -11 + 2.0 ** -14}


def _matrix(seed=21, n=30000):
    ip, cols, vals = synthetic.powerlaw_csr(n, seed, alpha=1.8, max_len=20000)
    vals = (vals * np.float32(3.0) ** synthetic.dense_uniform(seed + 1, vals.size) * 7).astype(np.float32)
    return ip, cols, vals


def test_bf16_planes_match_oracle(orc):
    ip, cols, vals = _matrix()
    n = ip.size - 1
    p = prepare_spmv_csr_split(to_dev(vals), to_dev(cols), to_dev(ip), (n, n), "bf16x2")
    hi, lo = orc.bf16_split(vals)
    assert np.array_equal(p.hi.cpu().numpy().view(np.uint16), hi)
    assert np.array_equal(p.lo.cpu().numpy().view(np.uint16), lo)
    assert np.array_equal(decoded_values(p).cpu().numpy(), orc.bf16_join(hi, lo))


def test_f16_rowscale_encoding():
    ip, cols, vals = _matrix(seed=5)
    n = ip.size - 1
    p = prepare_spmv_csr_split(to_dev(vals), to_dev(cols), to_dev(ip), (n, n), "f16_rowscale")
    scale = p.row_scale.cpu().numpy()
    lens = np.diff(ip)
    rmax = np.zeros(n, np.float32)
    np.maximum.at(rmax, np.repeat(np.arange(n), lens), np.abs(vals))
    nz = rmax > 0
    # power of two, and max|v| / s in (1/2, 1]
    assert np.all(np.log2(scale) == np.round(np.log2(scale)))
    assert np.all((rmax[nz] / scale[nz] > 0.5) & (rmax[nz] / scale[nz] <= 1.0))
    dec = decoded_values(p).cpu().numpy()
    s_of = np.repeat(scale, lens)
    assert np.all(np.abs(dec - vals) <= 2.0 ** -11 * np.abs(vals) + 2.0 ** -25 * s_of)


@pytest.mark.parametrize("fmt", ["bf16x2", "bf16", "f16_rowscale"])
@pytest.mark.parametrize("sch", [None, Schedule(group_lanes=4, short_max=64, chunk_nnz=256),
                                 Schedule(group_lanes=32, short_max=4096, chunk_nnz=4096, xcd_remap=1),
                                 Schedule(tile_nnz=2048, tile_mode=1),
                                 Schedule(tile_nnz=2048, tile_mode=1, x_slice_cols=1500)],
                         ids=["default", "s4", "s32", "entry", "sliced"])
def test_split_spmv_parity_and_stated_error(orc, fmt, sch):
    ip, cols, vals = _matrix()
    n = ip.size - 1
    x = synthetic.dense_uniform(33, n)
    p = prepare_spmv_csr_split(to_dev(vals), to_dev(cols), to_dev(ip), (n, n), fmt, schedule=sch)
    y = flagsparse_spmv_csr_split(to_dev(x), p)
    yh = host64(y)
    dec = decoded_values(p).cpu().numpy()
    # parity with the oracle on the stored (decoded) values
    y_dec = orc.spmv(ip, cols, dec, x)
    cond_dec = orc.spmv_abs(ip, cols, dec, x)
    assert row_rel_err(yh, y_dec, cond_dec).max() <= 1e-5
    # stated error against the original fp32 matrix
    y64 = orc.spmv(ip, cols, vals, x)
    cond = orc.spmv_abs(ip, cols, vals, x)
    assert row_rel_err(yh, y64, cond).max() <= BOUND_VS_FP32[fmt]
    # deterministic
    assert torch.equal(flagsparse_spmv_csr_split(to_dev(x), p), y)


def test_split_alpha_beta():
    ip, cols, vals = _matrix(n=5000)
    n = ip.size - 1
    x = to_dev(synthetic.dense_uniform(1, n))
    p = prepare_spmv_csr_split(to_dev(vals), to_dev(cols), to_dev(ip), (n, n), "bf16x2")
    base = host64(flagsparse_spmv_csr_split(x, p))
    y0 = synthetic.dense_uniform(2, n)
    out = to_dev(y0)
    res = flagsparse_spmv_csr_split(x, p, out=out, alpha=0.5, beta=3.0)
    assert res is out
    assert np.allclose(host64(res), 0.5 * base + 3.0 * y0, rtol=1e-5, atol=1e-5)


def test_split_rejects_non_fp32():
    d = torch.ones(2, dtype=torch.float64, device="cuda")
    with pytest.raises(TypeError):
        prepare_spmv_csr_split(d, torch.zeros(2, dtype=torch.int32, device="cuda"),
                               torch.tensor([0, 2], device="cuda"), (1, 1))
    with pytest.raises(ValueError):
        prepare_spmv_csr_split(d.float(), torch.zeros(2, dtype=torch.int32, device="cuda"),
                               torch.tensor([0, 2], device="cuda"), (1, 1), value_format="fp8")


@pytest.mark.slow
def test_bf16x2_powerlaw_4m_full_size(orc):
    """bf16 hi + lo storage on the 4M-row power-law matrix (48.6M entries) with
    the library's schedule: kernel parity on the decoded values (1e-5 of
    sum|a x| per row) and the stated bound against the fp32 matrix."""
    ip, cols, vals = synthetic.spmv_powerlaw_4m()
    n = ip.size - 1
    print(f"[bf16x2] powerlaw_4m nnz={vals.size}", flush=True)
    x = synthetic.dense_uniform(202, n)
    p = prepare_spmv_csr_split(to_dev(vals), to_dev(cols), to_dev(ip), (n, n), "bf16x2")
    y = host64(flagsparse_spmv_csr_split(to_dev(x), p))
    dec = decoded_values(p).cpu().numpy()
    y_dec = orc.spmv(ip, cols, dec, x)
    assert row_rel_err(y, y_dec, orc.spmv_abs(ip, cols, dec, x)).max() <= 1e-5
    y64 = orc.spmv(ip, cols, vals, x)
    cond = orc.spmv_abs(ip, cols, vals, x)
    assert row_rel_err(y, y64, cond).max() <= BOUND_VS_FP32["bf16x2"]


@pytest.mark.parametrize("fmt", ["bf16x2", "bf16"])
@pytest.mark.parametrize("sch", [
    None,                                                    # library schedule
    Schedule(tile_nnz=2048, tile_mode=1),                    # entry tiles, carries
    Schedule(tile_nnz=4096, tile_mode=1),                    # 8-byte plane loads
    Schedule(tile_nnz=2048, tile_mode=1, x_slice_cols=4096),  # sliced, compensated combine
    Schedule(group_lanes=8, short_max=2048, chunk_nnz=8192, unroll=4, tile_nnz=2048),
    Schedule(group_lanes=4, short_max=64, chunk_nnz=512, unroll=4),
    Schedule(tile_nnz=2048, tile_mode=1, x_window=1024),     # x windows, 9 of them
])
def test_split_long_rows_are_compensated(fmt, sch):
    """A 60,001-entry row 2^13, then 60,000 terms 2^-13 (x = 1), all exact in
    bf16: each small term is below half an ulp of the running fp32 sum, so
    plain fp32 accumulation returns 8192 (error 7.3, 9e-4 of sum|a x|).  The
    split modes sum in compensated fp32: the row comes out exact
    (8199.32421875), or within one rounding per column slice when sliced."""
    n_long = 60_001
    n = 9000
    lens = np.ones(n, np.int64)
    lens[7] = n_long
    ip = np.zeros(n + 1, np.int64)
    np.cumsum(lens, out=ip[1:])
    cols = np.concatenate([np.arange(lens[r]) % n for r in range(n)]).astype(np.int32)
    vals = np.full(ip[-1], 0.5, np.float32)
    vals[ip[7]] = 2.0 ** 13
    vals[ip[7] + 1:ip[8]] = 2.0 ** -13
    x = np.ones(n, np.float32)
    p = prepare_spmv_csr_split(to_dev(vals), to_dev(cols), to_dev(ip), (n, n), fmt,
                               schedule=sch)
    y = host64(flagsparse_spmv_csr_split(to_dev(x), p))
    # exact, up to one fp32 rounding of each column-slice partial (8 slices)
    assert abs(y[7] - (8192.0 + 60_000 * 2.0 ** -13)) <= 8 * 2.0 ** -11
    assert np.all(y[np.arange(n) != 7] == 0.5)


@pytest.mark.parametrize("T", [1024, 2048, 4096])
@pytest.mark.parametrize("fmt", ["bf16x2", "bf16", "f16_rowscale"])
def test_split_x_windows_equal_entry_tiles(orc, fmt, T):
    """Split storage over the x-window layout: the second pass reads the
    planes in CSR order with the gathered x (8-byte plane loads for full,
    aligned tiles), so with tiles of 2048 y equals the one-pass entry tiles
    of 2048 bit for bit; every tile size meets the decoded-values bar."""
    ip, cols, vals = _matrix(seed=41, n=60_000)
    n = ip.size - 1
    x = synthetic.dense_uniform(42, n)
    args = (to_dev(vals), to_dev(cols), to_dev(ip), (n, n), fmt)
    xw = prepare_spmv_csr_split(*args, schedule=Schedule(tile_nnz=T, tile_mode=1,
                                                         x_window=8192))
    y = flagsparse_spmv_csr_split(to_dev(x), xw)
    if T == 2048:
        ref = prepare_spmv_csr_split(*args, schedule=Schedule(tile_nnz=2048, tile_mode=1))
        assert torch.equal(y, flagsparse_spmv_csr_split(to_dev(x), ref))
    dec = decoded_values(xw).cpu().numpy()
    assert row_rel_err(host64(y), orc.spmv(ip, cols, dec, x),
                       orc.spmv_abs(ip, cols, dec, x)).max() <= 1e-5
