"""The uneven-batch workload: many independent square CSR matrices with n
log-uniform in [1e3, 2e5] and per-matrix Poisson row-length mean uniform in
[4, 64] (columns uniform, sorted per row, duplicates kept), fp32 values
U(-1, 1), int32 indices; each matrix gets an SpMV operand x and an
8-column SpMM operand B (both U(-1, 1)).  Generated on the device from a
seed; the default 512 matrices hold about 19M rows and 650M entries.
"""

from __future__ import annotations

import numpy as np
import torch


def uneven_batch(count=512, seed=4, device="cuda", n_dense=8, n_lo=1_000, n_hi=200_000,
                 mean_lo=4.0, mean_hi=64.0):
    """List of matrix dicts for BatchedCsr and a meta dict (rows, nnz)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ns = np.exp(rng.uniform(np.log(n_lo), np.log(n_hi), count)).round().astype(np.int64)
    means = rng.uniform(mean_lo, mean_hi, count)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    mats = []
    total_rows = total_nnz = 0
    for n, mean in zip(ns.tolist(), means.tolist()):
        lens = torch.poisson(torch.full((n,), mean, device=device), generator=gen).to(torch.int64)
        lens.clamp_(0, n)
        indptr = torch.zeros(n + 1, dtype=torch.int64, device=device)
        torch.cumsum(lens, 0, out=indptr[1:])
        nnz = int(indptr[-1].item())
        rows = torch.repeat_interleave(torch.arange(n, device=device), lens)
        key = rows * n + torch.randint(0, n, (nnz,), generator=gen, device=device)
        key, _ = torch.sort(key)
        cols = (key - rows * n).to(torch.int32)
        vals = torch.rand(nnz, generator=gen, device=device) * 2 - 1
        x = torch.rand(n, generator=gen, device=device) * 2 - 1
        B = torch.rand((n, n_dense), generator=gen, device=device) * 2 - 1
        mats.append(dict(data=vals, indices=cols, indptr=indptr.to(torch.int32), shape=(n, n),
                         x=x, B=B))
        total_rows += n
        total_nnz += nnz
        del rows, key
    return mats, {"count": count, "rows": total_rows, "nnz": total_nnz, "n_dense": n_dense}
