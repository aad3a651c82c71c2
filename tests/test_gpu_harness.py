"""The benchmark harness entry points (benchmarks.py:345-507, spmm_csr.py:2226-2593,
spmm_csr_opt_alg2.py:862-987): result layout, verification against torch
sparse, and the rocSPARSE vendor column in the reference's cusparse_* fields."""

import pytest
import torch

import flagsparse_public_amd as fs
from flagsparse_public_amd import vendor

pytestmark = pytest.mark.gpu

KEYS = {"parameters", "performance", "verification", "backend_status", "samples"}


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64, torch.complex64, torch.float16])
@pytest.mark.parametrize("op", ["non", "trans"])
def test_benchmark_spmv_case(dtype, op):
    r = fs.benchmark_spmv_case(n_rows=3000, n_cols=2000, nnz=40000, value_dtype=dtype,
                               warmup=2, iters=5, op=op)
    assert KEYS <= set(r)
    # the reference value is torch's own product in the value dtype (fp32 /
    # complex64 accumulate in fp32), so its strict allclose (atol 1e-6) can
    # miss near-zero rows for either side; the error itself is a few roundings
    assert isinstance(r["verification"]["triton_match_reference"], bool)
    bound = {torch.float32: 1e-4, torch.complex64: 1e-4, torch.float64: 1e-10,
             torch.float16: 5e-2}[dtype]
    assert r["verification"]["triton_max_error"] <= bound
    if dtype in (torch.float64, torch.float16):
        assert r["verification"]["triton_match_reference"]
    assert r["performance"]["triton_ms"] > 0
    if dtype == torch.float16:
        assert r["performance"]["cusparse_ms"] is None
        assert "float16" in r["backend_status"]["cusparse_unavailable_reason"]
    else:
        assert r["backend_status"]["cusparse_unavailable_reason"] is None, r["backend_status"]
        assert r["performance"]["cusparse_ms"] > 0
        # rocSPARSE accumulates fp32 in fp32, so its strict allclose against the
        # fp64-based reference (atol 1e-6) may miss near-zero rows; its error
        # stays at a few roundings of the row sums
        assert isinstance(r["verification"]["cusparse_match_reference"], bool)
        bound = 1e-10 if dtype == torch.float64 else 1e-4
        assert r["verification"]["cusparse_max_error"] <= bound


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_benchmark_spmm_cases(dtype):
    r = fs.benchmark_spmm_case(n_rows=2000, n_cols=1500, nnz=30000, n_dense_cols=16,
                               value_dtype=dtype, warmup=2, iters=5)
    assert KEYS <= set(r)
    assert r["verification"]["triton_strict_allclose_match"]
    # the vendor column ran and agrees within the reference's relative
    # threshold; its strict fp32 allclose (atol 1e-6) is rocSPARSE's own fp32
    # accumulation against an fp64 result and may miss near-zero entries
    v = r["verification"]
    assert v["cusparse_max_relative_error"] is not None, r["backend_status"]
    assert v["cusparse_max_relative_error"] <= v["cusparse_relative_threshold"]
    assert isinstance(v["cusparse_strict_allclose_match"], bool)
    r2 = fs.comprehensive_spmm_test(n_rows=500, n_cols=700, nnz=5000, n_dense_cols=8,
                                    dtype=dtype, warmup=1, iters=2)
    assert r2["parameters"]["n_dense_cols"] == 8
    r3 = fs.benchmark_spmm_opt_case(n_rows=2000, n_cols=1500, nnz=30000, n_dense_cols=32,
                                    value_dtype=dtype, warmup=2, iters=5)
    assert r3["verification"]["opt_ok"] and r3["performance"]["cu_ms"] > 0
    r4 = fs.benchmark_spmm_opt_alg2_case(n_rows=2000, n_cols=1500, nnz=30000, n_dense_cols=32,
                                         value_dtype=dtype, warmup=2, iters=5)
    v = r4["verification"]
    assert max(v["base_vs_torch_err"], v["opt_vs_torch_err"], v["opt_alg2_vs_torch_err"]) <= 1.0
    assert v["opt_vs_cusparse_err"] <= 1.0


def test_vendor_spmv_matches_library():
    """rocSPARSE adaptive and this library agree on a power-law matrix."""
    from flagsparse_public_amd import prepare_spmv_csr, flagsparse_spmv_csr, synthetic
    ip, cols, vals = synthetic.powerlaw_csr(20000, 3, alpha=1.7, max_len=8000)
    dev = "cuda"
    d = [torch.from_numpy(a).to(dev) for a in (vals, cols, ip.astype("int32"))]
    n = ip.size - 1
    x = torch.rand(n, device=dev) * 2 - 1
    y = torch.empty(n, device=dev)
    v = vendor.VendorSpmv(*d, (n, n), x, y, "adaptive")
    v.run()
    ours = flagsparse_spmv_csr(x=x, prepared=prepare_spmv_csr(*d, (n, n)), use_opt=True)
    assert torch.allclose(ours, y, rtol=1e-4, atol=1e-4)


def test_mtx_batch_harness(tmp_path):
    """The SuiteSparse harness over MatrixMarket files (tests/test_spmv.py:416-665):
    PASS rows for good files, an ERROR row for a broken one."""
    from flagsparse_public_amd import mtx
    import numpy as np
    rng = np.random.default_rng(5)
    paths = []
    for k, sym in enumerate(("general", "symmetric", "skew-symmetric")):
        n = 300 + 50 * k
        rows = rng.integers(0, n, 3000)
        cols = rng.integers(0, n, 3000)
        if sym != "general":
            rows, cols = np.maximum(rows, cols), np.minimum(rows, cols)
            keep = rows != cols if sym == "skew-symmetric" else np.ones(rows.size, bool)
            rows, cols = rows[keep], cols[keep]
        lines = [f"%%MatrixMarket matrix coordinate real {sym}", f"{n} {n} {rows.size}"]
        lines += [f"{r + 1} {c + 1} {float(v)!r}" for r, c, v in
                  zip(rows, cols, rng.uniform(-1, 1, rows.size))]
        p = tmp_path / f"m{k}.mtx"
        p.write_text("\n".join(lines) + "\n")
        paths.append(str(p))
    bad = tmp_path / "bad.mtx"
    bad.write_text("not a matrix\n")
    paths.append(str(bad))
    for op in ("non", "trans"):
        res = mtx.run_mtx_batch(paths, value_dtype=torch.float32, warmup=1, iters=2, op=op)
        assert [r["status"] for r in res] == ["PASS", "PASS", "PASS", "ERROR"], res
        assert all(r["err_cu"] is not None and r["err_cu"] <= 1.0 for r in res[:3])
    res = mtx.run_mtx_batch(paths[:1], value_dtype=torch.float64, warmup=1, iters=2)
    assert res[0]["status"] == "PASS" and res[0]["err_pt"] <= 1.0
