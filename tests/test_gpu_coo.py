"""COO SpMV (spmv_coo.py of the reference) against the fp64 oracle: dtypes,
index dtypes, op non / trans / conj, both sort_by_row settings, repeated
(row, col) pairs, empty inputs, and the validation contract."""

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from flagsparse_public_amd import PreparedCoo, flagsparse_spmv_coo, prepare_spmv_coo

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _coo(rng, n, m, k, npd, dup=True):
    r = rng.integers(0, n, k)
    c = rng.integers(0, m, k)
    if dup:  # repeated pairs
        r = np.concatenate([r, r[:50]])
        c = np.concatenate([c, c[:50]])
    v = rng.uniform(-1, 1, r.size)
    if np.dtype(npd).kind == "c":
        v = v + 1j * rng.uniform(-1, 1, r.size)
    return r, c, v.astype(npd)


@pytest.mark.parametrize("dtype,npd,tol", [(torch.float32, np.float32, 1e-5),
                                           (torch.float64, np.float64, 1e-12),
                                           (torch.complex64, np.complex64, 1e-5),
                                           (torch.complex128, np.complex128, 1e-12)],
                         ids=["f32", "f64", "c64", "c128"])
@pytest.mark.parametrize("idx", [torch.int32, torch.int64], ids=["i32", "i64"])
@pytest.mark.parametrize("op", ["non", "trans", "conj"])
@pytest.mark.parametrize("sort_by_row", [True, False])
def test_coo_spmv_matches_oracle(dtype, npd, tol, idx, op, sort_by_row):
    rng = np.random.default_rng(7)
    n, m = 700, 500
    r, c, v = _coo(rng, n, m, 6000, npd)
    A = sp.coo_matrix((v.astype(np.complex128 if dtype.is_complex else np.float64), (r, c)),
                      shape=(n, m)).tocsr()
    if op == "trans":
        A = A.T.tocsr()
    elif op == "conj":
        A = A.conj().T.tocsr()
    xlen = A.shape[1]
    x = rng.uniform(-1, 1, xlen)
    if dtype.is_complex:
        x = x + 1j * rng.uniform(-1, 1, xlen)
    x = x.astype(npd)
    y64 = A @ x.astype(np.complex128 if dtype.is_complex else np.float64)
    cond = abs(A) @ np.abs(x.astype(np.complex128 if dtype.is_complex else np.float64))
    args = (torch.from_numpy(v).to(DEV), torch.from_numpy(r).to(idx).to(DEV),
            torch.from_numpy(c).to(idx).to(DEV))
    y = flagsparse_spmv_coo(*args, torch.from_numpy(x).to(DEV), (n, m), op=op,
                            sort_by_row=sort_by_row)
    assert y.dtype == dtype and tuple(y.shape) == (A.shape[0],)
    got = y.cpu().numpy().astype(np.complex128 if dtype.is_complex else np.float64)
    assert (np.abs(got - y64) / np.maximum(cond, 1e-300)).max() <= tol
    p = prepare_spmv_coo(*args, (n, m), sort_by_row=sort_by_row, op=op)
    assert isinstance(p, PreparedCoo) and p.use_seg_kernel == sort_by_row
    y2 = flagsparse_spmv_coo(x=torch.from_numpy(x).to(DEV), prepared=p)
    assert torch.equal(y2, y)  # deterministic


def test_coo_spmv_empty_and_out():
    e = torch.empty(0, device=DEV)
    ei = torch.empty(0, dtype=torch.int32, device=DEV)
    y = flagsparse_spmv_coo(e, ei, ei, torch.ones(4, device=DEV), (3, 4))
    assert torch.equal(y, torch.zeros(3, device=DEV))
    v = torch.tensor([1.0, 2.0], device=DEV)
    r = torch.tensor([0, 2], dtype=torch.int32, device=DEV)
    c = torch.tensor([1, 0], dtype=torch.int32, device=DEV)
    out = torch.empty(3, device=DEV)
    res = flagsparse_spmv_coo(v, r, c, torch.tensor([3.0, 5.0], device=DEV), (3, 2), out=out)
    assert res is out and out.tolist() == [5.0, 0.0, 6.0]


def test_coo_spmv_validation():
    v = torch.tensor([1.0], device=DEV)
    i = torch.tensor([0], dtype=torch.int32, device=DEV)
    x = torch.ones(2, device=DEV)
    with pytest.raises(TypeError, match="COO SpMV supports"):
        flagsparse_spmv_coo(v.half(), i, i, x.half(), (2, 2))
    with pytest.raises(IndexError, match="row indices out of range"):
        flagsparse_spmv_coo(v, i + 5, i, x, (2, 2))
    with pytest.raises(IndexError, match="col indices out of range"):
        flagsparse_spmv_coo(v, i, i + 5, x, (2, 2))
    with pytest.raises(ValueError, match="same length"):
        flagsparse_spmv_coo(torch.ones(2, device=DEV), i, i, x, (2, 2))
    with pytest.raises(ValueError, match="op must be one of"):
        flagsparse_spmv_coo(v, i, i, x, (2, 2), op="bad")
    p = prepare_spmv_coo(v, i, i, (2, 2))
    with pytest.raises(ValueError, match="does not match prepared.shape"):
        flagsparse_spmv_coo(x=x, prepared=p, shape=(3, 2))
    with pytest.raises(ValueError, match="block_inner must be a positive power of 2"):
        flagsparse_spmv_coo(x=x, prepared=p, block_inner=3)
    with pytest.raises(ValueError, match="does not match prepared.op"):
        flagsparse_spmv_coo(x=x, prepared=p, op="trans")


@pytest.mark.parametrize("dtype,npd,tol", [(torch.float32, np.float32, 1e-6),
                                           (torch.float64, np.float64, 1e-12),
                                           (torch.complex64, np.complex64, 1e-6),
                                           (torch.float16, np.float16, 2e-3)],
                         ids=["f32", "f64", "c64", "f16"])
@pytest.mark.parametrize("idx", [torch.int32, torch.int64], ids=["i32", "i64"])
def test_coo_spmm_matches_oracle(dtype, npd, tol, idx):
    """flagsparse_spmm_coo (spmm_coo.py:671-695): widened accumulation, repeated
    pairs summed, against the fp64 product."""
    from flagsparse_public_amd import flagsparse_spmm_coo
    rng = np.random.default_rng(11)
    n, m, N = 400, 300, 24
    r, c, v = _coo(rng, n, m, 3000, npd)
    Bh = rng.uniform(-1, 1, (m, N))
    if dtype.is_complex:
        Bh = Bh + 1j * rng.uniform(-1, 1, (m, N))
    Bh = Bh.astype(npd)
    wide = np.complex128 if dtype.is_complex else np.float64
    A = sp.coo_matrix((v.astype(wide), (r, c)), shape=(n, m)).tocsr()
    C64 = A @ Bh.astype(wide)
    S = abs(A) @ np.abs(Bh.astype(wide))
    C = flagsparse_spmm_coo(torch.from_numpy(v).to(DEV), torch.from_numpy(r).to(idx).to(DEV),
                            torch.from_numpy(c).to(idx).to(DEV), torch.from_numpy(Bh).to(DEV),
                            (n, m))
    assert C.dtype == dtype and tuple(C.shape) == (n, N)
    got = C.cpu().numpy().astype(wide)
    assert (np.abs(got - C64) / np.maximum(S, 1e-300)).max() <= tol
    out = torch.empty((n, N), dtype=dtype, device=DEV)
    res, ms = flagsparse_spmm_coo(torch.from_numpy(v).to(DEV), torch.from_numpy(r).to(idx).to(DEV),
                                  torch.from_numpy(c).to(idx).to(DEV),
                                  torch.from_numpy(Bh).to(DEV), (n, m), out=out, return_time=True)
    assert res is out and ms >= 0 and torch.equal(out, C)


def test_coo_spmm_validation():
    from flagsparse_public_amd import flagsparse_spmm_coo
    v = torch.tensor([1.0], device=DEV)
    i = torch.tensor([0], dtype=torch.int32, device=DEV)
    B = torch.ones((2, 3), device=DEV)
    with pytest.raises(ValueError, match="B.shape\\[0\\] must be n_cols=3"):
        flagsparse_spmm_coo(v, i, i, B, (2, 3))
    with pytest.raises(TypeError, match="B dtype must match data dtype"):
        flagsparse_spmm_coo(v, i, i, B.double(), (2, 2))
    with pytest.raises(IndexError, match="row indices out of range for n_rows"):
        flagsparse_spmm_coo(v, i + 4, i, B, (2, 2))
    with pytest.raises(ValueError, match="block_n must be positive"):
        flagsparse_spmm_coo(v, i, i, B, (2, 2), block_n=0)
