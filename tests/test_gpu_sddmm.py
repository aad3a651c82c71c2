"""CSR SDDMM (sddmm_csr.py of the reference) against the fp64 oracle
(oracle.sddmm): K from 1 to 300 (vector and scalar loads, every group width),
alpha / beta, the diagnostic variants, empty cases and the validation
contract."""

import numpy as np
import pytest
import torch

from flagsparse_public_amd import flagsparse_sddmm_csr, prepare_sddmm_csr
from flagsparse_public_amd import synthetic

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.mark.parametrize("dtype,npd,tol", [(torch.float32, np.float32, 2e-6),
                                           (torch.float64, np.float64, 1e-13)],
                         ids=["f32", "f64"])
@pytest.mark.parametrize("k", [1, 3, 4, 16, 37, 64, 128, 300])
@pytest.mark.parametrize("with_data", [False, True])
def test_sddmm_matches_oracle(orc, dtype, npd, tol, k, with_data):
    ip, cols, _ = synthetic.powerlaw_csr(3000, 5, alpha=1.8, max_len=600)
    n = ip.size - 1
    rng = np.random.default_rng(k)
    x = rng.uniform(-1, 1, (n, k)).astype(npd)
    y = rng.uniform(-1, 1, (n, k)).astype(npd)
    data = rng.uniform(-1, 1, cols.size).astype(npd) if with_data else None
    alpha, beta = (1.5, -0.5) if with_data else (1.0, 0.0)
    want, cond = orc.sddmm(ip, cols, x, y, data, alpha, beta)
    out = flagsparse_sddmm_csr(None if data is None else torch.from_numpy(data).to(DEV),
                               torch.from_numpy(cols).to(DEV), torch.from_numpy(ip).to(DEV),
                               torch.from_numpy(x).to(DEV), torch.from_numpy(y).to(DEV), (n, n),
                               alpha=alpha, beta=beta)
    assert out.dtype == dtype and out.numel() == cols.size
    got = out.cpu().numpy().astype(np.float64)
    assert (np.abs(got - want) / np.maximum(cond, 1e-300)).max() <= tol


@pytest.mark.parametrize("nnz", [1, 5, 63, 65, 129, 1000])
@pytest.mark.parametrize("k", [4, 64, 300])
def test_sddmm_small_patterns_partial_blocks(orc, nnz, k):
    """Entry counts that leave the last workgroup partly empty (a lane group
    takes 2 / 4 / 8 entries for K = 4 / 64 / 300): every entry against the
    oracle, and nothing written past the last entry of the output."""
    rng = np.random.default_rng(nnz * 1000 + k)
    n_rows, n_cols = 13, 11
    rows = np.sort(rng.integers(0, n_rows, nnz))
    cols = rng.integers(0, n_cols, nnz).astype(np.int32)
    ip = np.zeros(n_rows + 1, np.int64)
    np.add.at(ip, rows + 1, 1)
    ip = np.cumsum(ip)
    x = rng.uniform(-1, 1, (n_rows, k)).astype(np.float32)
    y = rng.uniform(-1, 1, (n_cols, k)).astype(np.float32)
    want, cond = orc.sddmm(ip, cols, x, y, None, 1.0, 0.0)
    buf = torch.full((nnz + 512,), 7.0, device=DEV)
    out = flagsparse_sddmm_csr(indices=torch.from_numpy(cols).to(DEV),
                               indptr=torch.from_numpy(ip).to(DEV), x=torch.from_numpy(x).to(DEV),
                               y=torch.from_numpy(y).to(DEV), shape=(n_rows, n_cols),
                               out=buf[:nnz])
    got = out.cpu().numpy().astype(np.float64)
    assert (np.abs(got - want) / np.maximum(cond, 1e-300)).max() <= 2e-6
    assert torch.all(buf[nnz:] == 7.0)


def test_sddmm_variants_meta_and_prepared(orc):
    ip, cols, _ = synthetic.poisson_csr(2000, 10.0, 3)
    n = ip.size - 1
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, (n, 64)).astype(np.float32)
    y = rng.uniform(-1, 1, (n, 64)).astype(np.float32)
    want, cond = orc.sddmm(ip, cols, x, y)
    p = prepare_sddmm_csr(torch.from_numpy(cols).to(DEV), torch.from_numpy(ip).to(DEV), (n, n))
    xd, yd = torch.from_numpy(x).to(DEV), torch.from_numpy(y).to(DEV)
    out, ms, meta = flagsparse_sddmm_csr(x=xd, y=yd, prepared=p, return_time=True,
                                         return_meta=True)
    assert ms >= 0 and meta["fallback_used"] is False and meta["block_k"] == 64
    from flagsparse_public_amd.sparse_operations.sddmm_csr import _run_sddmm_prepared
    o64, m64 = _run_sddmm_prepared(p, xd, yd, None, 1.0, 0.0, None, variant="acc64_out64")
    assert o64.dtype == torch.float64 and m64["acc_dtype"] == "float64"
    # fp64 accumulation of fp32 products: within one fp64 rounding per term
    assert (np.abs(o64.cpu().numpy() - want) / np.maximum(cond, 1e-300)).max() <= 1e-12
    for variant in ("baseline", "altreduce", "acc64"):
        o, m = _run_sddmm_prepared(p, xd, yd, None, 1.0, 0.0, None, variant=variant)
        assert o.dtype == torch.float32 and m["variant"] == variant
        assert (np.abs(o.cpu().numpy() - want) / np.maximum(cond, 1e-300)).max() <= 2e-6


def test_sddmm_empty_and_validation():
    ip = torch.zeros(4, dtype=torch.int64, device=DEV)
    ix = torch.empty(0, dtype=torch.int32, device=DEV)
    x = torch.ones((3, 5), device=DEV)
    y = torch.ones((2, 5), device=DEV)
    out = flagsparse_sddmm_csr(indices=ix, indptr=ip, x=x, y=y, shape=(3, 2))
    assert out.numel() == 0
    ip2 = torch.tensor([0, 1], dtype=torch.int64, device=DEV)
    ix2 = torch.tensor([0], dtype=torch.int32, device=DEV)
    x1, y1 = torch.ones((1, 5), device=DEV), torch.ones((1, 5), device=DEV)
    with pytest.raises(TypeError, match="indices dtype must be torch.int32"):
        flagsparse_sddmm_csr(indices=ix2.long(), indptr=ip2, x=x1, y=y1, shape=(1, 1))
    with pytest.raises(ValueError, match="data is required when beta is non-zero"):
        flagsparse_sddmm_csr(indices=ix2, indptr=ip2, x=x1, y=y1, shape=(1, 1), beta=1.0)
    with pytest.raises(ValueError, match="same K dimension"):
        flagsparse_sddmm_csr(indices=ix2, indptr=ip2, x=x1, y=torch.ones((1, 4), device=DEV),
                             shape=(1, 1))
    with pytest.raises(IndexError, match="indices out of range"):
        flagsparse_sddmm_csr(indices=ix2 + 3, indptr=ip2, x=x1, y=y1, shape=(1, 1))
    # K = 0: beta * data
    d = torch.tensor([2.0], device=DEV)
    o = flagsparse_sddmm_csr(d, ix2, ip2, torch.ones((1, 0), device=DEV),
                             torch.ones((1, 0), device=DEV), (1, 1), beta=3.0)
    assert o.tolist() == [6.0]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_benchmark_sddmm_case_keys(dtype):
    from flagsparse_public_amd import benchmark_sddmm_case
    r = benchmark_sddmm_case(n_rows=512, n_cols=300, nnz=4000, k_dim=40, value_dtype=dtype,
                             warmup=2, iters=3, alpha=2.0, beta=0.5, run_cusparse=True)
    assert set(r) == {"parameters", "performance", "verification", "backend_status", "samples"}
    # the flag uses the reference's allclose (fp32 atol 1e-6), which two
    # summation orders of a 40-term dot can miss near zero; the error itself
    # stays at a few fp32 roundings of |x||y|
    assert isinstance(r["verification"]["triton_match_pytorch"], bool)
    bound = 1e-4 if dtype == torch.float32 else 1e-12
    assert r["verification"]["triton_max_abs_error"] <= bound
    assert r["performance"]["triton_ms"] > 0
    assert r["backend_status"]["cusparse_unavailable_reason"] is None, r["backend_status"]
    assert r["performance"]["cusparse_ms"] > 0
    # rocSPARSE's values agree with ours to a few roundings of the 40-term dot
    assert torch.allclose(r["samples"]["cusparse"], r["samples"]["triton"], rtol=1e-4,
                          atol=1e-4 if dtype == torch.float32 else 1e-10)


def test_sddmm_fp32_native_into_fp64_out(orc):
    """baseline variant with an fp64 out: fp32 accumulation stored widened."""
    from flagsparse_public_amd.sparse_operations.sddmm_csr import _run_sddmm_prepared
    ip, cols, _ = synthetic.poisson_csr(1000, 6.0, 5)
    n = ip.size - 1
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, (n, 24)).astype(np.float32)
    y = rng.uniform(-1, 1, (n, 24)).astype(np.float32)
    want, cond = orc.sddmm(ip, cols, x, y)
    p = prepare_sddmm_csr(torch.from_numpy(cols).to(DEV), torch.from_numpy(ip).to(DEV), (n, n))
    o, m = _run_sddmm_prepared(p, torch.from_numpy(x).to(DEV), torch.from_numpy(y).to(DEV), None,
                               1.0, 0.0, None, out_dtype=torch.float64)
    assert o.dtype == torch.float64 and m["acc_dtype"] == "float32"
    err = np.abs(o.cpu().numpy() - want) / np.maximum(cond, 1e-300)
    assert err.max() <= 2e-6 and err.max() > 0  # fp32 accumulation, not fp64


def test_sddmm_strided_out_and_prepared_reuse(orc):
    """A strided out view is filled in place (same object returned); one
    prepared pattern serves several x / y of different K."""
    ip, cols, _ = synthetic.poisson_csr(500, 7.0, 9)
    n = ip.size - 1
    p = prepare_sddmm_csr(torch.from_numpy(cols).to(DEV), torch.from_numpy(ip).to(DEV), (n, n))
    rng = np.random.default_rng(21)
    for k in (8, 33):
        x = rng.uniform(-1, 1, (n, k)).astype(np.float32)
        y = rng.uniform(-1, 1, (n, k)).astype(np.float32)
        want, cond = orc.sddmm(ip, cols, x, y)
        big = torch.zeros(2 * cols.size, device=DEV)
        view = big[::2]
        got = flagsparse_sddmm_csr(x=torch.from_numpy(x).to(DEV), y=torch.from_numpy(y).to(DEV),
                                   prepared=p, out=view)
        assert got is view
        assert torch.count_nonzero(big[1::2]) == 0  # only the view's elements written
        assert (np.abs(view.cpu().numpy() - want) / np.maximum(cond, 1e-300)).max() <= 2e-6
