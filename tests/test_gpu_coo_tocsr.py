"""COO -> CSR SpMV wrappers (spmv_csr.py:875-1016 of the reference):
``coo_to_csr_for_spmv``, ``prepare_spmv_coo_tocsr`` and
``flagsparse_spmv_coo_tocsr`` against the fp64 oracle.

Covers repeated (row, col) pairs, unsorted input, ``assume_sorted``, empty
input, ``out``, the prepared path and the validation contract.  The
conversion itself is pure index work and is checked bit for bit against a
numpy stable sort (on the CPU as well as on the device).
"""

import numpy as np
import pytest
import torch

from flagsparse_public_amd import flagsparse_spmv_coo_tocsr, prepare_spmv_coo_tocsr
from flagsparse_public_amd.sparse_operations.spmv_csr import coo_to_csr_for_spmv

from ._util import row_rel_err, ulp_distance_f32

DEV = "cuda"


def _coo(rng, n, m, k, npd, n_dup=80):
    r = rng.integers(0, n, k)
    c = rng.integers(0, m, k)
    # repeated pairs, placed after their first occurrence and in between
    pick = rng.integers(0, k, n_dup)
    r = np.concatenate([r, r[pick]])
    c = np.concatenate([c, c[pick]])
    v = rng.uniform(-1, 1, r.size).astype(npd)
    return r, c, v


def _host_csr(r, c, v, n, m):
    """Stable (row, col) sort on the host: repeated pairs keep input order."""
    order = np.argsort(r.astype(np.int64) * m + c, kind="stable")
    ip = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(r, minlength=n), out=ip[1:])
    return ip, c[order].astype(np.int64), v[order]


@pytest.mark.parametrize("idx", [torch.int32, torch.int64], ids=["i32", "i64"])
def test_conversion_is_a_stable_sort_cpu(idx):
    rng = np.random.default_rng(3)
    n, m = 300, 250
    r, c, v = _coo(rng, n, m, 4000, np.float32)
    ip, cols, vals = _host_csr(r, c, v, n, m)
    d, ci, p = coo_to_csr_for_spmv(torch.from_numpy(v), torch.from_numpy(r).to(idx),
                                   torch.from_numpy(c).to(idx), (n, m))
    assert ci.dtype == torch.int32 and p.dtype == torch.int64
    assert np.array_equal(p.numpy(), ip)
    assert np.array_equal(ci.numpy(), cols)
    assert np.array_equal(d.numpy(), vals)  # repeated pairs in input order


def test_conversion_sorted_input_untouched_cpu():
    rng = np.random.default_rng(4)
    n, m = 200, 200
    r, c, v = _coo(rng, n, m, 2000, np.float64, n_dup=0)
    ip, cols, vals = _host_csr(r, c, v, n, m)
    rr = np.repeat(np.arange(n), np.diff(ip))
    for assume in (False, True):
        d, ci, p = coo_to_csr_for_spmv(torch.from_numpy(vals), torch.from_numpy(rr),
                                       torch.from_numpy(cols), (n, m), assume_sorted=assume)
        assert np.array_equal(p.numpy(), ip) and np.array_equal(ci.numpy(), cols)
        assert np.array_equal(d.numpy(), vals)


def test_conversion_empty_cpu():
    d, ci, p = coo_to_csr_for_spmv(torch.zeros(0), torch.zeros(0, dtype=torch.int64),
                                   torch.zeros(0, dtype=torch.int64), (5, 7))
    assert d.numel() == 0 and ci.dtype == torch.int32 and p.tolist() == [0] * 6


@pytest.mark.gpu
@pytest.mark.parametrize("idx", [torch.int32, torch.int64], ids=["i32", "i64"])
@pytest.mark.parametrize("use_opt", [True, False], ids=["opt", "default"])
def test_coo_tocsr_spmv_f32_vs_oracle(orc, idx, use_opt):
    rng = np.random.default_rng(11)
    n, m = 900, 700
    r, c, v = _coo(rng, n, m, 9000, np.float32)
    ip, cols, vals = _host_csr(r, c, v, n, m)
    x = rng.uniform(-1, 1, m).astype(np.float32)
    y64 = orc.spmv(ip, cols, vals, x)
    cond = orc.spmv_abs(ip, cols, vals, x)
    args = (torch.from_numpy(v).to(DEV), torch.from_numpy(r).to(idx).to(DEV),
            torch.from_numpy(c).to(idx).to(DEV))
    xd = torch.from_numpy(x).to(DEV)
    y = flagsparse_spmv_coo_tocsr(*args, xd, (n, m), use_opt=use_opt)
    assert y.dtype == torch.float32 and tuple(y.shape) == (n,)
    yh = y.cpu().numpy()
    if use_opt:
        assert row_rel_err(yh.astype(np.float64), y64, cond).max() <= 1e-5
    else:  # fp64 accumulation rounded once: within 1 ulp of fl32(oracle)
        assert ulp_distance_f32(yh, y64.astype(np.float32)).max() <= 1
    # the prepared path gives the same bits
    p = prepare_spmv_coo_tocsr(*args, (n, m))
    y2 = flagsparse_spmv_coo_tocsr(x=xd, prepared=p, use_opt=use_opt)
    assert torch.equal(y, y2)
    # into `out`, returned as the same object
    out = torch.empty(n, dtype=torch.float32, device=DEV)
    y3 = flagsparse_spmv_coo_tocsr(x=xd, prepared=p, use_opt=use_opt, out=out)
    assert y3 is out and torch.equal(out, y)
    # the conversion on the device equals the host stable sort
    d, ci, pp = coo_to_csr_for_spmv(*args, (n, m))
    assert np.array_equal(pp.cpu().numpy(), ip)
    assert np.array_equal(ci.cpu().numpy(), cols)
    assert np.array_equal(d.cpu().numpy(), vals)


@pytest.mark.gpu
def test_coo_tocsr_repeated_pairs_summed_in_input_order(orc):
    """Three copies of one (row, col) pair whose fp32 sum depends on the
    order: a stable sort reproduces the input-order sum exactly."""
    v = np.array([1.0, 2.0 ** -24, -1.0, 2.0 ** -24, 0.5], np.float32)
    r = np.array([0, 0, 0, 1, 0])
    c = np.array([3, 3, 3, 0, 1])
    x = np.array([1.0, 1.0, 0.0, 1.0], np.float32)
    ip, cols, vals = _host_csr(r, c, v, 2, 4)
    assert list(cols) == [1, 3, 3, 3, 0]
    assert list(vals) == [0.5, 1.0, 2.0 ** -24, -1.0, 2.0 ** -24]
    args = [torch.from_numpy(a).to(DEV) for a in (v, r, c)]
    y = flagsparse_spmv_coo_tocsr(*args, torch.from_numpy(x).to(DEV), (2, 4), use_opt=False)
    assert y.cpu().numpy().tolist() == orc.spmv(ip, cols, vals, x).astype(np.float32).tolist()


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,npd,tol", [(torch.float64, np.float64, 1e-12),
                                           (torch.complex64, np.complex64, 1e-5)],
                         ids=["f64", "c64"])
def test_coo_tocsr_other_dtypes(orc, dtype, npd, tol):
    rng = np.random.default_rng(12)
    n, m = 400, 300
    r, c, v = _coo(rng, n, m, 3000, np.float64)
    if dtype.is_complex:
        v = v + 1j * rng.uniform(-1, 1, v.size)
    v = v.astype(npd)
    ip, cols, vals = _host_csr(r, c, v, n, m)
    x = rng.uniform(-1, 1, m)
    if dtype.is_complex:
        x = x + 1j * rng.uniform(-1, 1, m)
    x = x.astype(npd)
    y64 = orc.spmv(ip, cols, vals, x)
    cond = orc.spmv_abs(ip, cols, vals, x)
    y = flagsparse_spmv_coo_tocsr(*[torch.from_numpy(a).to(DEV) for a in (v, r, c, x)], (n, m))
    got = y.cpu().numpy().astype(np.complex128 if dtype.is_complex else np.float64)
    assert row_rel_err(got, y64, cond).max() <= tol


@pytest.mark.gpu
def test_coo_tocsr_assume_sorted_and_empty(orc):
    rng = np.random.default_rng(13)
    n, m = 500, 400
    r, c, v = _coo(rng, n, m, 4000, np.float32, n_dup=0)
    ip, cols, vals = _host_csr(r, c, v, n, m)
    rr = np.repeat(np.arange(n), np.diff(ip))
    x = rng.uniform(-1, 1, m).astype(np.float32)
    y64 = orc.spmv(ip, cols, vals, x)
    cond = orc.spmv_abs(ip, cols, vals, x)
    args = [torch.from_numpy(a).to(DEV) for a in (vals, rr, cols, x)]
    ya = flagsparse_spmv_coo_tocsr(*args, (n, m), assume_sorted=True)
    yb = flagsparse_spmv_coo_tocsr(*args, (n, m))
    assert torch.equal(ya, yb)
    assert row_rel_err(ya.double().cpu().numpy(), y64, cond).max() <= 1e-5
    # no entries: zeros of the right shape
    e = [torch.zeros(0, dtype=torch.float32, device=DEV),
         torch.zeros(0, dtype=torch.int64, device=DEV),
         torch.zeros(0, dtype=torch.int64, device=DEV)]
    ye = flagsparse_spmv_coo_tocsr(*e, torch.ones(6, device=DEV), (4, 6))
    assert ye.tolist() == [0.0] * 4
    pe = prepare_spmv_coo_tocsr(*e, (4, 6))
    assert flagsparse_spmv_coo_tocsr(x=torch.ones(6, device=DEV), prepared=pe).tolist() == [0.0] * 4


@pytest.mark.gpu
def test_coo_tocsr_validation():
    v = torch.ones(3, device=DEV)
    r = torch.tensor([0, 1, 1], device=DEV)
    c = torch.tensor([0, 0, 2], device=DEV)
    x = torch.ones(3, device=DEV)
    with pytest.raises(TypeError, match="must all be torch.Tensor"):
        flagsparse_spmv_coo_tocsr(v.tolist(), r, c, x, (2, 3))
    with pytest.raises(ValueError, match="CUDA tensors"):
        flagsparse_spmv_coo_tocsr(v.cpu(), r, c, x, (2, 3))
    with pytest.raises(ValueError, match="1D"):
        flagsparse_spmv_coo_tocsr(v[None], r, c, x, (2, 3))
    with pytest.raises(TypeError, match="x dtype must match"):
        flagsparse_spmv_coo_tocsr(v, r, c, x.double(), (2, 3))
    with pytest.raises(TypeError, match="data dtype must be one of"):
        flagsparse_spmv_coo_tocsr(v.int(), r, c, x.int(), (2, 3))
    with pytest.raises(ValueError, match="same length"):
        prepare_spmv_coo_tocsr(v, r[:2], c[:2], (2, 3))
    p = prepare_spmv_coo_tocsr(v, r, c, (2, 3))
    with pytest.raises(TypeError, match="x is required"):
        flagsparse_spmv_coo_tocsr(prepared=p)
    with pytest.raises(ValueError, match="does not match prepared.shape"):
        flagsparse_spmv_coo_tocsr(x=x, prepared=p, shape=(3, 3))
    with pytest.raises(IndexError):
        flagsparse_spmv_coo_tocsr(v, r, torch.tensor([0, 0, 5], device=DEV), x, (2, 3))
