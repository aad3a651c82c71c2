"""GPU parity of the HIP CSR SpMM against the CPU oracle (MI355X).

Bars per output element against the fp64 oracle C64 and the condition
S = |A| |B|: fp32 default path (fp64 accumulation): within 1 fp32 ulp of
fl32(C64) everywhere, identical on >= 99.9%; fp32 opt / alg2 (native):
|C - C64| <= 1e-5 S; fp64: 1e-12 S; fp16 / bf16: the reference suite's
allclose (5e-3 / 1e-1); complex: 1e-5 / 1e-12.
"""

import numpy as np
import pytest
import torch

from tests._util import golden_case, host64, to_dev, ulp_distance_f32

pytestmark = pytest.mark.gpu

from flagsparse_public_amd import (  # noqa: E402
    flagsparse_spmm_csr,
    flagsparse_spmm_csr_opt,
    flagsparse_spmm_csr_opt_alg2,
    prepare_spmm_csr_opt,
    prepare_spmm_csr_opt_alg2,
    synthetic,
)
from flagsparse_public_amd.rowsplit import Schedule  # noqa: E402

DEV = "cuda"


def _check(C, C64, S, dtype, path):
    if C.is_complex():
        got = C.cpu().to(torch.complex128).numpy()
    else:
        got = host64(C)
    if dtype == torch.float32 and path == "default":
        ulps = ulp_distance_f32(got.astype(np.float32), C64.astype(np.float32))
        assert ulps.max() <= 1
        assert (ulps == 0).mean() >= 0.999
        return
    rel = np.abs(got - C64) / np.where(S > 0, S, 1.0)
    if dtype in (torch.float32, torch.complex64):
        assert rel.max() <= 1e-5
    elif dtype in (torch.float64, torch.complex128):
        assert rel.max() <= 1e-12
    elif dtype == torch.float16:
        assert np.allclose(got, C64, rtol=5e-3, atol=5e-3)
    else:
        assert np.allclose(got, C64, rtol=1e-1, atol=1e-1)


def _rand_csr(rng, n, m, mean, dtype):
    lens = np.minimum(rng.poisson(mean, n), m)
    ip = np.zeros(n + 1, np.int64)
    np.cumsum(lens, out=ip[1:])
    cols = rng.integers(0, m, ip[-1]).astype(np.int32)  # duplicates, unsorted
    if np.dtype(dtype).kind == "c":
        vals = (rng.normal(size=ip[-1]) + 1j * rng.normal(size=ip[-1])).astype(dtype)
    else:
        vals = rng.normal(size=ip[-1]).astype(dtype)
    return ip, cols, vals


def _dense(rng, shape, dtype):
    if np.dtype(dtype).kind == "c":
        return (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(dtype)
    return rng.normal(size=shape).astype(dtype)


@pytest.mark.parametrize("name", ["ka2x2", "longrow", "unsorted_dup", "empty_rows", "all_empty", "cancel"])
def test_spmm_golden_exact(name):
    c = golden_case(name)
    shape = tuple(int(v) for v in c["shape"])
    C = flagsparse_spmm_csr(to_dev(c["data"]), to_dev(c["indices"].astype(np.int32)),
                            to_dev(c["indptr"]), to_dev(c["B"]), shape)
    if name == "cancel":
        # exact result 0: fp64 accumulation (like the reference's default path)
        # leaves a residue bounded by (len + 1) u64 sum|a b|
        S = np.abs(c["data"]).astype(np.float64)[:, None] * np.abs(c["B"][c["indices"]]).astype(np.float64)
        bound = 14 * 2.0 ** -53 * S.sum(axis=0)
        assert np.all(np.abs(C.cpu().numpy()[0].astype(np.float64) - c["C_exact"][0]) <= bound)
        assert np.array_equal(C.cpu().numpy()[1], c["C_f32"][1])
    elif c["data"].dtype == np.float32:
        assert np.array_equal(C.cpu().numpy(), c["C_f32"])
    else:
        assert np.allclose(C.cpu().numpy(), c["C_exact"], rtol=1e-15, atol=1e-300)


NP = {torch.float16: np.float16, torch.bfloat16: np.float32, torch.float32: np.float32,
      torch.float64: np.float64, torch.complex64: np.complex64, torch.complex128: np.complex128}


@pytest.mark.parametrize("dtype", list(NP), ids=lambda d: str(d).replace("torch.", ""))
@pytest.mark.parametrize("n_dense", [1, 4, 5, 12, 24, 32, 48, 64, 96, 128])
def test_spmm_default_path_grid(orc, dtype, n_dense):
    rng = np.random.default_rng(n_dense)
    ip, cols, vals = _rand_csr(rng, 300, 200, 6.0, NP[dtype])
    B = _dense(rng, (200, n_dense), NP[dtype])
    data = to_dev(vals, dtype)
    Bd = to_dev(B, dtype)
    # oracle on the values actually stored (fp16 / bf16 rounding included)
    v64 = data.cpu().to(torch.complex128 if dtype.is_complex else torch.float64).numpy()
    B64 = Bd.cpu().to(torch.complex128 if dtype.is_complex else torch.float64).numpy()
    C64 = orc.spmm(ip, cols, v64, B64)
    S = orc.spmm_abs(ip, cols, v64, B64)
    C = flagsparse_spmm_csr(data, to_dev(cols), to_dev(ip), Bd, (300, 200))
    assert C.dtype == dtype and tuple(C.shape) == (300, n_dense)
    _check(C, C64, S, dtype, "default")


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("n_dense", [1, 8, 32, 64, 96, 128, 200])
@pytest.mark.parametrize("sch", [None, Schedule(short_max=16, chunk_nnz=64),
                                 Schedule(short_max=0, chunk_nnz=128, xcd_remap=1),
                                 Schedule(short_max=64, chunk_nnz=256, col_tile=32),
                                 Schedule(short_max=16, chunk_nnz=64, unroll=8),
                                 Schedule(short_max=64, chunk_nnz=256, col_tile=64, unroll=8)],
                         ids=["default", "chunky", "allchunk_xcd", "tile32", "chunky_u8",
                              "tile64_u8"])
def test_spmm_opt_and_alg2_schedules(orc, dtype, n_dense, sch):
    rng = np.random.default_rng(100 + n_dense)
    npd = NP[dtype]
    lens = np.concatenate([rng.poisson(5, 400), [0, 3000, 700, 129, 2049]])
    ip = np.zeros(lens.size + 1, np.int64)
    np.cumsum(lens, out=ip[1:])
    m = 3500
    cols = rng.integers(0, m, ip[-1]).astype(np.int32)
    vals = rng.uniform(-1, 1, ip[-1]).astype(npd)
    B = rng.uniform(-1, 1, (m, n_dense)).astype(npd)
    C64 = orc.spmm(ip, cols, vals, B)
    S = orc.spmm_abs(ip, cols, vals, B)
    shape = (lens.size, m)
    p = prepare_spmm_csr_opt(to_dev(vals), to_dev(cols), to_dev(ip), shape, schedule=sch)
    Bd = to_dev(B)
    C1 = flagsparse_spmm_csr_opt(B=Bd, prepared=p)
    _check(C1, C64, S, dtype, "opt")
    for _ in range(2):  # deterministic chunk fix-up
        assert torch.equal(flagsparse_spmm_csr_opt(B=Bd, prepared=p), C1)
    p2 = prepare_spmm_csr_opt_alg2(to_dev(vals), to_dev(cols), to_dev(ip), shape, schedule=sch)
    C2, meta = flagsparse_spmm_csr_opt_alg2(B=Bd, prepared=p2, return_meta=True)
    _check(C2, C64, S, dtype, "opt")
    assert meta["warp_size"] == 64 and meta["n_dense_cols"] == n_dense


def test_spmm_alpha_beta_noncontiguous_out():
    rng = np.random.default_rng(9)
    ip, cols, vals = _rand_csr(rng, 500, 400, 7.0, np.float32)
    B = rng.uniform(-1, 1, (16, 400)).astype(np.float32)
    Bt = to_dev(B).t()  # non-contiguous B
    assert not Bt.is_contiguous()
    args = (to_dev(vals), to_dev(cols), to_dev(ip))
    base = flagsparse_spmm_csr(*args, Bt, (500, 400))
    C0 = rng.uniform(-1, 1, (500, 16)).astype(np.float32)
    out = to_dev(C0)
    res = flagsparse_spmm_csr(*args, Bt, (500, 400), out=out, alpha=-1.5, beta=2.0)
    assert res is out
    assert np.allclose(host64(res), -1.5 * host64(base) + 2.0 * C0, rtol=1e-5, atol=1e-5)


def test_spmm_beta_with_strided_out():
    """A strided out (a transposed view and a column slice) with beta != 0:
    C = alpha A B + beta out, on the ALG1, opt and alg2 paths."""
    rng = np.random.default_rng(19)
    ip, cols, vals = _rand_csr(rng, 300, 260, 6.0, np.float32)
    B = to_dev(rng.uniform(-1, 1, (260, 12)).astype(np.float32))
    args = (to_dev(vals), to_dev(cols), to_dev(ip))
    base = host64(flagsparse_spmm_csr(*args, B, (300, 260)))
    C0 = rng.uniform(-1, 1, (300, 12)).astype(np.float32)
    pm = prepare_spmm_csr_opt(*args, (300, 260))
    p2 = prepare_spmm_csr_opt_alg2(*args, (300, 260))
    runs = [lambda out: flagsparse_spmm_csr(*args, B, (300, 260), out=out, alpha=0.5, beta=-2.0),
            lambda out: flagsparse_spmm_csr_opt(B=B, prepared=pm, out=out, alpha=0.5, beta=-2.0),
            lambda out: flagsparse_spmm_csr_opt_alg2(B=B, prepared=p2, out=out, alpha=0.5,
                                                     beta=-2.0)]
    for run in runs:
        holder = to_dev(np.ascontiguousarray(C0.T))   # (12, 300)
        out = holder.t()                                # (300, 12), strided
        assert not out.is_contiguous()
        assert run(out) is out
        assert np.allclose(host64(out), 0.5 * base - 2.0 * C0, rtol=1e-5, atol=1e-5)
        wide = torch.full((300, 20), 9.0, device=DEV)
        wide[:, 4:16] = to_dev(C0)
        out = wide[:, 4:16]
        assert not out.is_contiguous()
        assert run(out) is out
        assert np.allclose(host64(out), 0.5 * base - 2.0 * C0, rtol=1e-5, atol=1e-5)
        assert torch.all(wide[:, :4] == 9.0) and torch.all(wide[:, 16:] == 9.0)


def test_spmm_api_validation():
    """The reference's negative API cases (tests/test_spmm.py:728-748)."""
    d = torch.tensor([1.0, 2.0, 3.0], device=DEV)
    i = torch.tensor([0, 1, 1], dtype=torch.int32, device=DEV)
    p = torch.tensor([0, 2, 3], dtype=torch.int64, device=DEV)
    B = torch.randn((2, 4), device=DEV)
    long_cols = torch.cat([torch.arange(1536), torch.tensor([7, 129, 511, 1024])]).to(torch.int32).to(DEV)
    long_d = torch.randn(long_cols.numel(), device=DEV)
    long_p = torch.tensor([0, 1536, 1540], dtype=torch.int64, device=DEV)
    long_B = torch.randn((2048, 32), device=DEV)
    cases = [
        (lambda: flagsparse_spmm_csr(d, i, p, B, (2,)), ValueError),
        (lambda: flagsparse_spmm_csr(d, i, p, B[0], (2, 2)), ValueError),
        (lambda: flagsparse_spmm_csr(d, i, p, B.double(), (2, 2)), TypeError),
        (lambda: flagsparse_spmm_csr(d, i, p, torch.randn((3, 4), device=DEV), (2, 2)), ValueError),
        (lambda: flagsparse_spmm_csr(d, i, p[:-1], B, (2, 2)), ValueError),
        (lambda: flagsparse_spmm_csr(d, i, torch.tensor([1, 2, 3], device=DEV), B, (2, 2)), ValueError),
        (lambda: flagsparse_spmm_csr(d, i, torch.tensor([0, 2, 2], device=DEV), B, (2, 2)), ValueError),
        (lambda: flagsparse_spmm_csr(d, i, torch.tensor([0, 3, 2], device=DEV), B, (2, 2)), ValueError),
        (lambda: flagsparse_spmm_csr(d, torch.tensor([0, 3, 1], dtype=torch.int32, device=DEV), p, B, (2, 2)), IndexError),
        (lambda: flagsparse_spmm_csr(d, i, p, B, (2, 2), block_n=0), ValueError),
        (lambda: flagsparse_spmm_csr(d, i, p, B, (2, 2), block_nnz=0), ValueError),
        (lambda: flagsparse_spmm_csr(d, i, p, B, (2, 2), max_segments=0), ValueError),
        (lambda: flagsparse_spmm_csr(d, i, p, B, (2, 2), out=torch.empty((3, 4), device=DEV)), ValueError),
        (lambda: flagsparse_spmm_csr(d, i, p, B, (2, 2), out=torch.empty((2, 4))), ValueError),
        (lambda: flagsparse_spmm_csr(long_d, long_cols, long_p, long_B, (2, 2048), block_nnz=128,
                                     max_segments=4), ValueError),
    ]
    for fn, exc in cases:
        with pytest.raises(exc):
            fn()
    out = torch.empty((2, 4), device=DEV)
    assert flagsparse_spmm_csr(d, i, p, B, (2, 2), out=out) is out
    C = flagsparse_spmm_csr(torch.empty(0, device=DEV), torch.empty(0, dtype=torch.int32, device=DEV),
                            torch.zeros(3, dtype=torch.int64, device=DEV), B, (2, 2))
    assert torch.equal(C, torch.zeros(2, 4, device=DEV))
    assert flagsparse_spmm_csr(d, i, p, torch.empty((2, 0), device=DEV), (2, 2)).shape == (2, 0)


def test_alg2_prepared_out_and_meta():
    rng = np.random.default_rng(3)
    ip, cols, vals = _rand_csr(rng, 16, 32, 6.0, np.float32)
    B = rng.normal(size=(32, 24)).astype(np.float32)
    p = prepare_spmm_csr_opt_alg2(to_dev(vals), to_dev(cols).long(), to_dev(ip), (16, 32))
    out_buf = torch.empty((16, 24), device=DEV)
    out, ms, meta = flagsparse_spmm_csr_opt_alg2(B=to_dev(B), prepared=p, out=out_buf,
                                                 return_time=True, return_meta=True)
    assert out is out_buf and isinstance(ms, float)
    assert meta["device_name"] and isinstance(meta["sm_count"], int)
    assert meta["max_row_nnz"] == p.max_row_nnz and meta["bucket_hits"] and meta["launch_configs"]


@pytest.mark.slow
@pytest.mark.parametrize("n_dense,dtype", [(32, torch.float32), (64, torch.float32),
                                           (128, torch.float32), (64, torch.float64)])
def test_spmm_banded_1m_workload(orc, n_dense, dtype):
    """1,048,576-row banded + random SpMM at full size."""
    npd = NP[dtype]
    ip, cols, vals = synthetic.spmm_banded_1m(dtype=npd)
    n = ip.size - 1
    B = synthetic.dense_uniform(303, (n, n_dense), npd)
    C64 = orc.spmm(ip, cols, vals, B)
    S = orc.spmm_abs(ip, cols, vals, B)
    p = prepare_spmm_csr_opt(to_dev(vals), to_dev(cols), to_dev(ip), (n, n))
    _check(flagsparse_spmm_csr_opt(B=to_dev(B), prepared=p), C64, S, dtype, "opt")


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
def test_spmm_diagnose_candidate(orc, dtype):
    """The diagnose hooks (spmm_csr.py:1835-2031): candidate rows labelled by
    bucket, and the candidate C accumulated in fp64 -- for fp32 within one
    ulp of the rounded fp64 oracle and bit-identical on nearly every entry."""
    from flagsparse_public_amd.sparse_operations import spmm_csr as m
    rng = np.random.default_rng(12)
    npd = NP[dtype]
    lens = np.concatenate([rng.poisson(20, 600), [0, 3000, 2100, 700, 129]])
    ip = np.zeros(lens.size + 1, np.int64)
    np.cumsum(lens, out=ip[1:])
    m_cols = 4000
    cols = rng.integers(0, m_cols, ip[-1]).astype(np.int32)
    vals = rng.uniform(-1, 1, ip[-1]).astype(npd)
    B = rng.uniform(-1, 1, (m_cols, 24)).astype(npd)
    C64 = orc.spmm(ip, cols, vals, B)
    S = orc.spmm_abs(ip, cols, vals, B)
    p = prepare_spmm_csr_opt(to_dev(vals), to_dev(cols), to_dev(ip), (lens.size, m_cols))
    labels = {b["label"] for b in m._build_spmm_opt_candidate_buckets(p)}
    assert "split" in labels and "short_16" in labels
    C = m._flagsparse_spmm_csr_opt_candidate_for_diagnose(p, to_dev(B))
    _check(C, C64, S, dtype, "default")
