"""Helpers shared by the tests: numpy <-> torch CSR, error measures."""

import os

import numpy as np
import torch

GOLDEN_PATH = os.path.join(os.path.dirname(__file__), "golden", "csr_golden.npz")

NP_OF = {
    torch.float16: np.float16,
    torch.bfloat16: np.float32,  # numpy has no bf16; values are widened
    torch.float32: np.float32,
    torch.float64: np.float64,
    torch.complex64: np.complex64,
    torch.complex128: np.complex128,
}


def golden_case(name):
    g = np.load(GOLDEN_PATH)
    pre = name + "__"
    return {k[len(pre):]: g[k] for k in g.files if k.startswith(pre)}


def to_dev(arr, dtype=None, device="cuda"):
    t = torch.from_numpy(np.ascontiguousarray(arr))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(device)


def host64(t):
    """Device tensor -> numpy fp64 / complex128 (bf16 / fp16 widened)."""
    t = t.detach().cpu()
    if t.is_complex():
        return t.to(torch.complex128).numpy()
    return t.to(torch.float64).numpy()


def random_mask_csr(M, N, dtype, index_dtype, seed, lo=0.06, hi=0.25, scale=32.0):
    """Random-mask CSR like the reference's accuracy suite
    (tests/pytest/test_spmv_csr_accuracy.py:69-81): density
    clamp(scale / (M N), lo, hi), sorted unique columns, normal values."""
    g = torch.Generator().manual_seed(seed)
    p = min(hi, max(lo, scale / max(M * N, 1)))
    mask = torch.rand(M, N, generator=g) < p
    if int(mask.sum()) == 0:
        mask[0, 0] = True
    if dtype.is_complex:
        comp = torch.float32 if dtype == torch.complex64 else torch.float64
        dense = torch.complex(torch.randn(M, N, generator=g, dtype=comp),
                              torch.randn(M, N, generator=g, dtype=comp))
    else:
        dense = torch.randn(M, N, generator=g, dtype=torch.float32).to(dtype)
    dense = torch.where(mask, dense, torch.zeros((), dtype=dtype))
    rows, cols = torch.nonzero(mask, as_tuple=True)
    data = dense[rows, cols].contiguous()
    indptr = torch.zeros(M + 1, dtype=torch.int64)
    indptr[1:] = torch.cumsum(torch.bincount(rows, minlength=M), 0)
    return data, cols.to(index_dtype), indptr.to(index_dtype), dense


def row_rel_err(y, y64, cond):
    """|y - y64| / sum|a x| per row (0 where the row is empty)."""
    d = np.abs(np.asarray(y) - y64)
    return np.where(cond > 0, d / np.where(cond > 0, cond, 1.0), d)


def ulp_distance_f32(a, b):
    ai = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    bi = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    ai = np.where(ai < 0, np.int64(-2**31) - ai, ai)
    bi = np.where(bi < 0, np.int64(-2**31) - bi, bi)
    return np.abs(ai - bi)
