"""GPU parity of the HIP CSR SpMV against the CPU oracle (MI355X).

Bars (per row, against the fp64 oracle y64 and the condition sum|a x|):
  * fp32 default path (fp64 accumulation, one rounding): bit-identical to
    fl32(y64) on >= 99.9% of rows and within 1 fp32 ulp on all rows;
  * fp32 opt path (native accumulation): |y - y64| <= 1e-5 * sum|a x|;
  * fp64: |y - y64| <= 1e-12 * sum|a x|;
  * complex64: 1e-5, complex128: 1e-12 (same measure on |.|);
  * fp16 / bf16: the reference suite's allclose (5e-3 / 1e-1).
"""

import importlib

import numpy as np
import pytest
import torch

from tests._util import golden_case, host64, random_mask_csr, row_rel_err, to_dev, ulp_distance_f32

pytestmark = pytest.mark.gpu

spmv_mod = importlib.import_module("flagsparse_public_amd.sparse_operations.spmv_csr")
from flagsparse_public_amd import flagsparse_spmv_csr, prepare_spmv_csr  # noqa: E402
from flagsparse_public_amd.rowsplit import Schedule  # noqa: E402
from flagsparse_public_amd import synthetic  # noqa: E402

DEV = "cuda"


def _oracle_parts(orc, data, indices, indptr, x):
    d = host64(data) if not data.is_complex() else data.cpu().to(torch.complex128).numpy()
    xx = host64(x) if not x.is_complex() else x.cpu().to(torch.complex128).numpy()
    ip = indptr.cpu().numpy().astype(np.int64)
    ix = indices.cpu().numpy().astype(np.int64)
    return orc.spmv(ip, ix, d, xx), orc.spmv_abs(ip, ix, d, xx)


def _check(y, y64, cond, dtype, path):
    y = host64(y) if not y.is_complex() else y.cpu().to(torch.complex128).numpy()
    if dtype == torch.float32 and path == "default":
        ulps = ulp_distance_f32(y.astype(np.float32), y64.astype(np.float32))
        assert ulps.max() <= 1, f"max ulp {ulps.max()}"
        assert (ulps == 0).mean() >= 0.999
    elif dtype in (torch.float32, torch.complex64):
        assert row_rel_err(y, y64, cond).max() <= 1e-5
    elif dtype in (torch.float64, torch.complex128):
        assert row_rel_err(y, y64, cond).max() <= 1e-12
    elif dtype == torch.float16:
        assert np.allclose(y, y64, rtol=5e-3, atol=5e-3)
    else:
        assert np.allclose(y, y64, rtol=1e-1, atol=1e-1)


@pytest.mark.parametrize("name", ["ka2x2", "longrow", "unsorted_dup", "empty_rows", "all_empty",
                                  "cancel", "complex"])
def test_spmv_golden_exact(name):
    c = golden_case(name)
    data = to_dev(c["data"])
    indices = to_dev(c["indices"].astype(np.int32))
    indptr = to_dev(c["indptr"])
    x = to_dev(c["x"])
    shape = tuple(int(v) for v in c["shape"])
    y = flagsparse_spmv_csr(data, indices, indptr, x, shape)
    if data.dtype == torch.float32:
        # default path: fp64 accumulation rounded once == exactly rounded result
        assert np.array_equal(y.cpu().numpy(), c["y_f32"])
    else:
        got = y.cpu().numpy()
        cond = np.abs(c["y_exact"]) + 1e-300
        assert np.all(np.abs(got - c["y_exact"]) <= 1e-15 * np.maximum(cond, 1.0))


@pytest.mark.parametrize("M,N", [(1, 32), (160, 1024), (128, 256)])
@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16, torch.float32, torch.float64,
                                   torch.complex64, torch.complex128],
                         ids=["float16", "bfloat16", "float32", "float64", "complex64", "complex128"])
@pytest.mark.parametrize("index_dtype", [torch.int32, torch.int64], ids=["int32", "int64"])
@pytest.mark.parametrize("op", ["non", "trans", "conj"])
def test_spmv_matches_oracle_grid(orc, M, N, dtype, index_dtype, op):
    """The reference accuracy grid (test_spmv_csr_accuracy.py:108-132), checked
    against the fp64 oracle instead of a dense torch product."""
    seed = M * 7919 + N
    data, indices, indptr, dense = random_mask_csr(M, N, dtype, index_dtype, seed)
    transpose = op in ("trans", "conj")
    g = torch.Generator().manual_seed(seed + 1)
    xlen = M if transpose else N
    if dtype.is_complex:
        comp = torch.float32 if dtype == torch.complex64 else torch.float64
        x = torch.complex(torch.randn(xlen, generator=g, dtype=comp),
                          torch.randn(xlen, generator=g, dtype=comp))
    else:
        x = torch.randn(xlen, generator=g).to(dtype)
    ref_dense = dense.to(torch.complex128 if dtype.is_complex else torch.float64)
    if op == "trans":
        ref_dense = ref_dense.t()
    elif op == "conj":
        ref_dense = ref_dense.conj().t().resolve_conj()
    A = ref_dense.numpy()
    xx = x.to(torch.complex128 if dtype.is_complex else torch.float64).numpy()
    y64 = A @ xx
    cond = np.abs(A) @ np.abs(xx)
    y = flagsparse_spmv_csr(data.to(DEV), indices.to(DEV), indptr.to(DEV), x.to(DEV), (M, N), op=op)
    assert y.dtype == dtype and tuple(y.shape) == (N if transpose else M,)
    _check(y, y64, cond, dtype, "default")
    if dtype in (torch.float32, torch.float64) and index_dtype == torch.int32:
        y2 = flagsparse_spmv_csr(data.to(DEV), indices.to(DEV), indptr.to(DEV), x.to(DEV), (M, N),
                                 op=op, use_opt=True)
        _check(y2, y64, cond, dtype, "opt")


def _powerlaw(n, seed, max_len, dtype=np.float32):
    return synthetic.powerlaw_csr(n, seed, alpha=1.6, max_len=max_len, dtype=dtype)


SCHEDULES = [
    Schedule(group_lanes=s, short_max=sm, chunk_nnz=ch, unroll=u, xcd_remap=xr)
    for s, sm, ch, u, xr in [
        (1, 16, 64, 4, 0), (2, 32, 128, 2, 1), (4, 64, 512, 4, 0), (8, 128, 512, 8, 1),
        (16, 256, 1024, 4, 0), (32, 512, 2048, 4, 1), (64, 4096, 4096, 8, 0), (8, 0, 64, 4, 0),
    ]
] + [
    Schedule(tile_nnz=t, short_max=sm, chunk_nnz=ch, unroll=u, xcd_remap=xr)
    for t, sm, ch, u, xr in [
        (256, 64, 512, 4, 0), (1024, 1024, 2048, 8, 1), (2048, 2048, 4096, 2, 0),
        (2048, 256, 8192, 8, 1), (512, 0, 64, 4, 0),
    ]
] + [
    # entry tiles: fixed runs of entries, rows cut at tile boundaries
    Schedule(tile_nnz=t, tile_mode=1) for t in (1024, 2048, 4096)
] + [
    # column slices (native path): 7 and 3 slices of the 20000 columns
    Schedule(tile_nnz=1024, short_max=1024, chunk_nnz=2048, unroll=4, x_slice_cols=3000),
    Schedule(group_lanes=8, short_max=128, chunk_nnz=512, unroll=4, x_slice_cols=7000),
]


@pytest.mark.parametrize("sch", SCHEDULES, ids=lambda s: f"S{s.group_lanes}_t{s.tile_nnz}_m{s.short_max}_c{s.chunk_nnz}_u{s.unroll}_x{s.xcd_remap}_sl{s.x_slice_cols}_tm{s.tile_mode}")
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("ptr_dtype", [torch.int32, torch.int64], ids=["p32", "p64"])
def test_spmv_schedules_powerlaw(orc, sch, dtype, ptr_dtype):
    npd = np.float32 if dtype == torch.float32 else np.float64
    ip, cols, vals = _powerlaw(20000, 11, 30000, npd)
    x = synthetic.dense_uniform(12, ip.size - 1, npd)
    y64 = orc.spmv(ip, cols, vals, x)
    cond = orc.spmv_abs(ip, cols, vals, x)
    data, indices, indptr = to_dev(vals), to_dev(cols), to_dev(ip).to(ptr_dtype)
    xd = to_dev(x)
    n = ip.size - 1
    p = prepare_spmv_csr(data, indices, indptr, (n, n), schedule=sch)
    y_def = flagsparse_spmv_csr(x=xd, prepared=p)
    _check(y_def, y64, cond, dtype, "default")
    y_opt = flagsparse_spmv_csr(x=xd, prepared=p, use_opt=True)
    _check(y_opt, y64, cond, dtype, "opt")
    # deterministic: the chunked rows' fix-up sums partials in chunk order
    for _ in range(3):
        assert torch.equal(flagsparse_spmv_csr(x=xd, prepared=p, use_opt=True), y_opt)
        assert torch.equal(flagsparse_spmv_csr(x=xd, prepared=p), y_def)


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16, torch.float32, torch.float64,
                                   torch.complex64, torch.complex128],
                         ids=["f16", "bf16", "f32", "f64", "c64", "c128"])
@pytest.mark.parametrize("index_dtype", [torch.int32, torch.int64], ids=["i32", "i64"])
def test_spmv_entry_tiles_all_dtypes(orc, dtype, index_dtype):
    """Entry tiles (2048 entries; 1024 and 4096 where built) on a power-law
    matrix with rows far longer than a tile, for every value / index dtype,
    both accumulation paths, alpha / beta, and misaligned arrays (scalar
    loads instead of 16-byte ones)."""
    npd = {torch.float16: np.float32, torch.bfloat16: np.float32, torch.float32: np.float32,
           torch.float64: np.float64, torch.complex64: np.complex64,
           torch.complex128: np.complex128}[dtype]
    ip, cols, vals = _powerlaw(9000, 23, 20000, np.float64)
    n = ip.size - 1
    rng = np.random.default_rng(31)
    x = rng.uniform(-1, 1, n)
    if dtype.is_complex:
        vals = vals + 1j * rng.uniform(-1, 1, vals.size)
        x = x + 1j * rng.uniform(-1, 1, n)
    vals = vals.astype(npd)
    x = x.astype(npd)
    data = to_dev(vals).to(dtype)
    xd = to_dev(x).to(dtype)
    # oracle on the values as stored (fp16 / bf16 rounding included)
    vs = data.cpu().to(torch.complex128 if dtype.is_complex else torch.float64).numpy()
    xs = xd.cpu().to(torch.complex128 if dtype.is_complex else torch.float64).numpy()
    y64 = orc.spmv(ip, cols.astype(np.int64), vs, xs)
    cond = orc.spmv_abs(ip, cols.astype(np.int64), vs, xs)
    hot = dtype in (torch.float32, torch.float64) and index_dtype == torch.int32
    sizes = (1024, 2048, 4096) if hot else (2048,)
    for T in sizes:
        for misalign in (False, True):
            if misalign:  # one element of padding in front: no 16-byte loads
                d2 = torch.empty(data.numel() + 1, dtype=dtype, device=DEV)[1:]
                d2.copy_(data)
                c2 = torch.empty(cols.size + 1, dtype=index_dtype, device=DEV)[1:]
                c2.copy_(to_dev(cols).to(index_dtype))
            else:
                d2, c2 = data, to_dev(cols).to(index_dtype)
            sch = Schedule(tile_nnz=T, tile_mode=1)
            p = prepare_spmv_csr(d2, c2, to_dev(ip), (n, n), schedule=sch)
            for use_opt in ((False, True) if dtype in (torch.float32, torch.float64) else (False,)):
                y = flagsparse_spmv_csr(x=xd, prepared=p, use_opt=use_opt)
                got = y.cpu().to(torch.complex128 if dtype.is_complex else torch.float64).numpy()
                err = row_rel_err(got, y64, cond).max()
                bound = {torch.float16: 1e-3, torch.bfloat16: 8e-3, torch.float32: 1e-6,
                         torch.float64: 1e-13, torch.complex64: 1e-6, torch.complex128: 1e-13}[dtype]
                assert err <= bound, (T, misalign, use_opt, err)
            # y = 2 A x - 0.5 y_old
            y0 = to_dev(rng.uniform(-1, 1, n).astype(npd)).to(dtype)
            y1 = y0.clone()
            spmv_mod._hip_spmv_csr_impl_prepared(p, xd, out=y1, alpha=2.0, beta=-0.5)
            want = 2.0 * y64 - 0.5 * y0.cpu().to(torch.complex128 if dtype.is_complex else torch.float64).numpy()  # noqa: E501
            got = y1.cpu().to(torch.complex128 if dtype.is_complex else torch.float64).numpy()
            y0h = y0.cpu().to(torch.complex128 if dtype.is_complex else torch.float64).numpy()
            scale_ = 2.0 * cond + 0.5 * np.abs(y0h)
            assert (np.abs(got - want) / np.maximum(scale_, 1e-300)).max() <= bound


@pytest.mark.parametrize("n_rows", [30001, 29900], ids=["118blocks", "117blocks"])
@pytest.mark.parametrize("use_opt", [False, True], ids=["sums", "native"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64, torch.complex64],
                         ids=["f32", "f64", "c64"])
def test_sliced_combine_row_masks_equal_row_ids(dtype, use_opt, n_rows):
    """The combine's two ways to find a row's partials -- presence masks with
    popcounts, or the compact-row ids -- give the same bits (fp64 row sums of
    the default fp32 path, native partials otherwise), including 24 slices
    (three groups of 8), a row count that ends inside a 32-row mask word, an
    odd number of 256-row blocks (the mask kernel takes two per workgroup) and
    alpha / beta."""
    from flagsparse_public_amd import _lib
    npd = {torch.float32: np.float32, torch.float64: np.float64,
           torch.complex64: np.complex64}[dtype]
    ip, cols, vals = _powerlaw(n_rows, 5, 20000, np.float64)
    n = ip.size - 1
    assert n % 32 != 0
    rng = np.random.default_rng(8)
    if dtype.is_complex:
        vals = vals + 1j * rng.uniform(-1, 1, vals.size)
    vals = vals.astype(npd)
    x = rng.uniform(-1, 1, n).astype(npd)
    for slice_cols in (4000, 1251):
        sch = Schedule(tile_nnz=2048, tile_mode=1, x_slice_cols=slice_cols)
        p = prepare_spmv_csr(to_dev(vals), to_dev(cols), to_dev(ip), (n, n), schedule=sch)
        flagsparse_spmv_csr(x=to_dev(x), prepared=p, use_opt=use_opt)  # fills the partials
        sliced = [v for k, v in p._plans.items() if isinstance(k, tuple) and k[0] == "sliced"]
        assert sliced
        s = sliced[-1]
        assert s.n_slices == (8 if slice_cols == 4000 else 24)
        assert s.desc.row_mask
        no_mask = _lib.FsSlices(**{f: getattr(s.desc, f) for f, _ in _lib.FsSlices._fields_})
        no_mask.row_mask = None
        y0 = to_dev(rng.uniform(-1, 1, n).astype(npd))
        y1, y2 = y0.clone(), y0.clone()
        _lib.sliced_combine(s.desc, s.z, y1, 1.5, -0.25, y1.device)
        _lib.sliced_combine(no_mask, s.z, y2, 1.5, -0.25, y2.device)
        assert torch.equal(y1, y2), slice_cols


@pytest.mark.parametrize("index_dtype", [np.int32, np.int64], ids=["i32", "i64"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("pattern", ["banded", "random"])
def test_fixed_rule_large_x(orc, pattern, dtype, index_dtype):
    """The fixed rule on a matrix whose x exceeds 8 MB (2.5M columns): a
    sample of rows finds banded columns local (plain entry tiles of 2048)
    and random columns not (x-window tiles of 4096, any column type); both
    paths match the oracle."""
    rng = np.random.default_rng(21)
    n = 2_500_000
    if pattern == "banded":
        off = np.arange(-2, 3)
        ip = np.arange(0, 5 * n + 1, 5, dtype=np.int64)
        cols = np.clip(np.repeat(np.arange(n), 5) + np.tile(off, n), 0, n - 1)
    else:
        lens = rng.integers(0, 8, n)
        ip = np.zeros(n + 1, np.int64)
        np.cumsum(lens, out=ip[1:])
        cols = rng.integers(0, n, int(ip[-1]))
    cols = cols.astype(index_dtype)
    npd = np.float32 if dtype == torch.float32 else np.float64
    vals = rng.uniform(-1, 1, cols.size).astype(npd)
    x = rng.uniform(-1, 1, n).astype(npd)
    y64, cond = orc.spmv(ip, cols, vals, x), orc.spmv_abs(ip, cols, vals, x)
    p = prepare_spmv_csr(to_dev(vals), to_dev(cols), to_dev(ip), (n, n), schedule="default")
    want = (Schedule(tile_nnz=2048, tile_mode=1) if pattern == "banded"
            else Schedule(tile_nnz=4096, tile_mode=1, x_window=-1))
    assert p.schedule == want
    xd = to_dev(x)
    _check(flagsparse_spmv_csr(x=xd, prepared=p, use_opt=True), y64, cond, dtype, "opt")
    _check(flagsparse_spmv_csr(x=xd, prepared=p), y64, cond, dtype, "default")


def test_spmv_long_rows_compensated_fp32(orc):
    """Rows far above 4096 entries with heavy cancellation stay within 1e-5 of
    sum|a x| on the native fp32 path (compensated chunk accumulation)."""
    rng = np.random.default_rng(5)
    n, m = 8, 300000
    lens = np.array([250000, 100000, 5000, 4097, 1, 0, 60000, 3])
    ip = np.zeros(n + 1, np.int64)
    np.cumsum(lens, out=ip[1:])
    cols = np.concatenate([np.sort(rng.choice(m, l, replace=False)) for l in lens]).astype(np.int32)
    vals = (rng.uniform(-1, 1, ip[-1]) * 1e3).astype(np.float32)
    x = rng.uniform(-1, 1, m).astype(np.float32)
    y64 = orc.spmv(ip, cols, vals, x)
    cond = orc.spmv_abs(ip, cols, vals, x)
    y = flagsparse_spmv_csr(to_dev(vals), to_dev(cols), to_dev(ip), to_dev(x), (n, m), use_opt=True)
    assert row_rel_err(host64(y), y64, cond).max() <= 1e-6


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64, torch.complex64], ids=["f32", "f64", "c64"])
def test_spmv_tiles_many_empty_rows(orc, dtype):
    """Tiles are capped at 512 rows: a matrix of mostly empty rows with a few
    dense and long ones."""
    rng = np.random.default_rng(17)
    n, m = 70000, 12000
    lens = np.zeros(n, np.int64)
    hot = rng.choice(n, 300, replace=False)
    lens[hot] = rng.integers(1, 40, 300)
    lens[hot[:5]] = [3000, 4500, 2049, 2048, 9000]
    ip = np.zeros(n + 1, np.int64)
    np.cumsum(lens, out=ip[1:])
    cols = np.concatenate([np.sort(rng.choice(m, l, replace=False)) for l in lens if l]).astype(np.int32)
    npd = {torch.float32: np.float32, torch.float64: np.float64, torch.complex64: np.complex64}[dtype]
    vals = rng.uniform(-1, 1, ip[-1]).astype(npd)
    x = rng.uniform(-1, 1, m).astype(npd)
    if dtype.is_complex:
        vals = (vals + 1j * rng.uniform(-1, 1, ip[-1])).astype(npd)
        x = (x + 1j * rng.uniform(-1, 1, m)).astype(npd)
    y64 = orc.spmv(ip, cols, vals, x)
    cond = orc.spmv_abs(ip, cols, vals, x)
    big = 1024 if dtype.is_complex else 2048  # complex row tiles: 16-byte sums in LDS
    for sch in (Schedule(tile_nnz=big, short_max=big, chunk_nnz=4096, unroll=8),
                Schedule(tile_nnz=256, short_max=256, chunk_nnz=256, unroll=4),
                Schedule(tile_nnz=2048, tile_mode=1)):
        p = prepare_spmv_csr(to_dev(vals), to_dev(cols), to_dev(ip), (n, m), schedule=sch)
        for use_opt in (False, True):
            y = flagsparse_spmv_csr(x=to_dev(x), prepared=p, use_opt=use_opt)
            got = y.cpu().to(torch.complex128).numpy() if dtype.is_complex else host64(y)
            tol = 1e-12 if dtype == torch.float64 else 1e-6
            assert row_rel_err(got, y64, cond).max() <= tol


def test_spmv_alpha_beta_and_out():
    ip, cols, vals = synthetic.poisson_csr(3000, 9.0, 3, lo=0)
    x = synthetic.dense_uniform(4, 3000)
    y0 = synthetic.dense_uniform(5, 3000)
    data, indices, indptr, xd = to_dev(vals), to_dev(cols), to_dev(ip), to_dev(x)
    base = flagsparse_spmv_csr(data, indices, indptr, xd, (3000, 3000), use_opt=True)
    out = to_dev(y0)
    res = flagsparse_spmv_csr(data, indices, indptr, xd, (3000, 3000), out=out, alpha=2.0,
                              beta=-0.5, use_opt=True)
    assert res is out
    expect = 2.0 * host64(base) - 0.5 * y0.astype(np.float64)
    assert np.allclose(host64(res), expect, rtol=1e-6, atol=1e-6)
    out2 = torch.full((3000,), float("nan"), device=DEV)
    res2 = flagsparse_spmv_csr(data, indices, indptr, xd, (3000, 3000), out=out2)
    assert res2 is out2 and not torch.isnan(res2).any()


@pytest.mark.parametrize("use_opt", [True, False], ids=["opt", "default"])
def test_spmv_alpha_without_out_and_strided_out(use_opt):
    """alpha applies with out=None; a strided out gets alpha A x + beta out
    (staged through a contiguous copy) and the entries between its elements
    stay untouched."""
    n = 2500
    ip, cols, vals = synthetic.poisson_csr(n, 9.0, 13, lo=0)
    x = synthetic.dense_uniform(14, n)
    y0 = synthetic.dense_uniform(15, n)
    data, indices, indptr, xd = to_dev(vals), to_dev(cols), to_dev(ip), to_dev(x)
    p = prepare_spmv_csr(data, indices, indptr, (n, n))
    base = host64(flagsparse_spmv_csr(x=xd, prepared=p, use_opt=use_opt))
    y = flagsparse_spmv_csr(x=xd, prepared=p, use_opt=use_opt, alpha=2.0)
    assert np.allclose(host64(y), 2.0 * base, rtol=1e-6, atol=1e-6)
    y = flagsparse_spmv_csr(data, indices, indptr, xd, (n, n), use_opt=use_opt, alpha=-3.0)
    assert np.allclose(host64(y), -3.0 * base, rtol=1e-6, atol=1e-6)
    buf = torch.full((2 * n,), 7.0, device=DEV)
    buf[::2] = to_dev(y0)
    out = buf[::2]
    assert not out.is_contiguous()
    res = flagsparse_spmv_csr(x=xd, prepared=p, use_opt=use_opt, out=out, alpha=2.0, beta=-0.5)
    assert res is out
    assert np.allclose(host64(res), 2.0 * base - 0.5 * y0.astype(np.float64), rtol=1e-6,
                       atol=1e-6)
    assert torch.all(buf[1::2] == 7.0)
    # beta = 0: the strided out is written, not read
    buf[::2] = float("nan")
    res = flagsparse_spmv_csr(x=xd, prepared=p, use_opt=use_opt, out=out, alpha=0.5)
    assert np.allclose(host64(res), 0.5 * base, rtol=1e-6, atol=1e-6)


def test_spmv_empty_shapes():
    e = torch.empty(0, device=DEV)
    y = flagsparse_spmv_csr(e, torch.empty(0, dtype=torch.int32, device=DEV),
                            torch.zeros(1, dtype=torch.int64, device=DEV),
                            torch.ones(5, device=DEV), (0, 5))
    assert y.shape == (0,)
    y = flagsparse_spmv_csr(e, torch.empty(0, dtype=torch.int32, device=DEV),
                            torch.zeros(4, dtype=torch.int64, device=DEV),
                            torch.ones(5, device=DEV), (3, 5))
    assert torch.equal(y, torch.zeros(3, device=DEV))


def test_spmv_16k_workload(orc):
    """16384-row Poisson(16) SpMV, all paths."""
    ip, cols, vals = synthetic.spmv_16k()
    x = synthetic.dense_uniform(101, 16384)
    y64 = orc.spmv(ip, cols, vals, x)
    cond = orc.spmv_abs(ip, cols, vals, x)
    args = (to_dev(vals), to_dev(cols), to_dev(ip), to_dev(x), (16384, 16384))
    _check(flagsparse_spmv_csr(*args), y64, cond, torch.float32, "default")
    _check(flagsparse_spmv_csr(*args, use_opt=True), y64, cond, torch.float32, "opt")


@pytest.mark.slow
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
def test_spmv_powerlaw_4m_workload(orc, dtype):
    """The 4M-row power-law SpMV at full size against the oracle."""
    npd = np.float32 if dtype == torch.float32 else np.float64
    ip, cols, vals = synthetic.spmv_powerlaw_4m(dtype=npd)
    n = ip.size - 1
    x = synthetic.dense_uniform(202, n, npd)
    y64 = orc.spmv(ip, cols, vals, x)
    cond = orc.spmv_abs(ip, cols, vals, x)
    p = prepare_spmv_csr(to_dev(vals), to_dev(cols), to_dev(ip), (n, n))
    xd = to_dev(x)
    _check(flagsparse_spmv_csr(x=xd, prepared=p, use_opt=True), y64, cond, dtype, "opt")
    if dtype == torch.float32:
        _check(flagsparse_spmv_csr(x=xd, prepared=p), y64, cond, dtype, "default")


# ---- the reference's own API / fallback tests (test_spmv_csr_accuracy.py:135-232)

def test_prepared_transpose_mismatch_rejected():
    data, indices, indptr, _ = random_mask_csr(8, 10, torch.float32, torch.int32, 1)
    p = spmv_mod.prepare_spmv_csr(data.to(DEV), indices.to(DEV), indptr.to(DEV), (8, 10),
                                  transpose=True)
    with pytest.raises(ValueError, match="does not match prepared.transpose"):
        flagsparse_spmv_csr(x=torch.randn(8, device=DEV), prepared=p, transpose=False)


def test_prepared_op_mismatch_rejected():
    data, indices, indptr, _ = random_mask_csr(8, 10, torch.complex64, torch.int32, 2)
    p = spmv_mod.prepare_spmv_csr(data.to(DEV), indices.to(DEV), indptr.to(DEV), (8, 10), op="conj")
    x = torch.randn(8, dtype=torch.complex64, device=DEV)
    with pytest.raises(ValueError, match="does not match prepared.op"):
        flagsparse_spmv_csr(x=x, prepared=p, op="trans")


def test_int64_auto_fallback_to_int32(monkeypatch):
    data, indices, indptr, dense = random_mask_csr(12, 9, torch.float32, torch.int64, 3)
    x = torch.randn(9)
    ref = dense.double() @ x.double()
    state = {"forced_once": False}
    original = spmv_mod._triton_spmv_csr_impl_prepared

    def fail_int64_once(prepared, x_in):
        if prepared.kernel_indices.dtype == torch.int64 and not state["forced_once"]:
            state["forced_once"] = True
            raise RuntimeError("forced int64 launch failure")
        return original(prepared, x_in)

    monkeypatch.setattr(spmv_mod, "_triton_spmv_csr_impl_prepared", fail_int64_once)
    out = flagsparse_spmv_csr(data.to(DEV), indices.to(DEV), indptr.to(DEV), x.to(DEV), (12, 9),
                              index_fallback_policy="auto")
    assert state["forced_once"]
    assert torch.allclose(out.double().cpu(), ref, rtol=1e-4, atol=1e-4)


def test_int64_strict_no_fallback(monkeypatch):
    data, indices, indptr, _ = random_mask_csr(12, 9, torch.float32, torch.int64, 4)

    def fail_int64(prepared, x_in):
        if prepared.kernel_indices.dtype == torch.int64:
            raise RuntimeError("forced int64 launch failure")
        return spmv_mod._hip_spmv_csr_impl_prepared(prepared, x_in)

    monkeypatch.setattr(spmv_mod, "_triton_spmv_csr_impl_prepared", fail_int64)
    with pytest.raises(RuntimeError, match="forced int64 launch failure"):
        flagsparse_spmv_csr(data.to(DEV), indices.to(DEV), indptr.to(DEV),
                            torch.randn(9, device=DEV), (12, 9), index_fallback_policy="strict")


def test_spmv_validation_errors():
    d = torch.tensor([1.0, 2.0, 3.0], device=DEV)
    i = torch.tensor([0, 1, 1], dtype=torch.int32, device=DEV)
    p = torch.tensor([0, 2, 3], dtype=torch.int64, device=DEV)
    x = torch.ones(2, device=DEV)
    with pytest.raises(ValueError, match="indptr must start at zero"):
        flagsparse_spmv_csr(d, i, torch.tensor([1, 2, 3], device=DEV), x, (2, 2))
    with pytest.raises(ValueError, match=r"indptr\[-1\] must equal nnz"):
        flagsparse_spmv_csr(d, i, torch.tensor([0, 2, 2], device=DEV), x, (2, 2))
    # the reference checks indptr[-1] == nnz before monotonicity
    with pytest.raises(ValueError, match=r"indptr\[-1\] must equal nnz"):
        flagsparse_spmv_csr(d, i, torch.tensor([0, 3, 2], device=DEV), x, (2, 2))
    with pytest.raises(ValueError, match="non-decreasing"):
        flagsparse_spmv_csr(d, i, torch.tensor([0, 3, 1, 3], device=DEV), x, (3, 2))
    with pytest.raises(IndexError):
        flagsparse_spmv_csr(d, torch.tensor([0, 3, 1], dtype=torch.int32, device=DEV), p, x, (2, 2))
    with pytest.raises(TypeError, match="x dtype"):
        flagsparse_spmv_csr(d, i, p, x.double(), (2, 2))
    with pytest.raises(ValueError, match="x length"):
        flagsparse_spmv_csr(d, i, p, torch.ones(3, device=DEV), (2, 2))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_exhaustive_schedule_with_int64_columns(orc, dtype):
    """Exhaustive selection over all variants on an int64-column matrix with
    the strict index policy: variants whose one-pass tiles exist only for int32
    columns run as the 2048-entry tiles, so nothing fails, and the chosen
    schedule meets the oracle's bar."""
    npd = np.float32 if dtype == torch.float32 else np.float64
    ip, cols, vals = synthetic.powerlaw_csr(30_000, 41, alpha=1.9, max_len=5000, dtype=npd,
                                            index_dtype=np.int64)
    n = ip.size - 1
    x = synthetic.dense_uniform(42, n, dtype=npd)
    args = [to_dev(vals), to_dev(cols), to_dev(ip.astype(np.int64))]
    p = prepare_spmv_csr(*args, (n, n), schedule="exhaustive", index_fallback_policy="strict")
    assert p.kernel_indices.dtype == torch.int64
    sch = p.schedule
    assert not (sch.tile_mode == 1 and sch.x_window == 0 and sch.tile_nnz not in (0, 2048))
    y = flagsparse_spmv_csr(x=to_dev(x), prepared=p, use_opt=True)
    y64 = orc.spmv(ip, cols, vals, x)
    cond = orc.spmv_abs(ip, cols, vals, x)
    # int64 columns take the default path (fp64 accumulation for fp32), as in the reference
    assert row_rel_err(host64(y), y64, cond).max() <= (1e-6 if dtype == torch.float32 else 1e-12)


def test_stream_handle_follows_current_stream():
    """The library launches on torch's current stream: the raw handle it
    passes through the C-ABI is the one torch.cuda.stream selects, for a
    device given as torch.device, "cuda" or an index, and a prepared SpMV
    issued on a side stream (default priority) equals the one on the default
    stream, bit for bit."""
    from flagsparse_public_amd import _lib
    dev = torch.device("cuda", 0)
    cur = torch.cuda.current_stream(dev).cuda_stream
    assert _lib.stream_handle(dev) == cur
    assert _lib.stream_handle("cuda") == cur and _lib.stream_handle(0) == cur
    side = torch.cuda.Stream(device=dev)
    with torch.cuda.stream(side):
        assert _lib.stream_handle(dev) == side.cuda_stream != cur
    ip, cols, vals = synthetic.spmv_16k()
    n = ip.size - 1
    p = prepare_spmv_csr(*(torch.from_numpy(a).to(dev) for a in (vals, cols, ip)), (n, n))
    x = torch.from_numpy(synthetic.dense_uniform(11, n)).to(dev)
    y0 = flagsparse_spmv_csr(x=x, prepared=p, use_opt=True)
    side.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(side):
        y1 = flagsparse_spmv_csr(x=x, prepared=p, use_opt=True)
    torch.cuda.current_stream(dev).wait_stream(side)
    torch.cuda.synchronize()
    assert torch.equal(y0, y1)
