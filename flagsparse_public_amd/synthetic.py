"""Seeded synthetic CSR workloads (numpy PCG64 on the host).

Each generator returns host numpy arrays ``(indptr int64, indices, values)``
with sorted, duplicate-free column indices per row unless stated.  Sizes and
distributions:

* ``poisson_csr``   -- square n, row lengths clip(Poisson(mean), lo, n),
  columns uniform without replacement (the 16384-row SpMV workload).
* ``powerlaw_csr``  -- row lengths ceil(l) with l drawn from the continuous
  density ~ l^-alpha on [1, max_len] (alpha = 2: mean ~ ln(max_len) + 0.58,
  about 12.1 for max_len = 1e5, i.e. ~48.4M entries at n = 4M; with the
  ceiling almost every row has >= 2 entries), columns uniform without
  replacement (the 4M-row SpMV workload).
* ``banded_random_csr`` -- offsets -h..h clipped at the edges plus k uniform
  off-band distinct columns per row (the 1,048,576-row SpMM workload:
  h = 3, k = 8, at most 15 entries per row).
* ``block_diag_csr``, ``rmat_csr``, ``banded_csr`` -- the remaining selector
  families.

Values are U(-1, 1) in the requested dtype.
"""

from __future__ import annotations

import math

import numpy as np


def _rng(seed) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def _values(rng, n, dtype):
    dt = np.dtype(dtype)
    if dt.kind == "c":
        comp = np.float32 if dt == np.complex64 else np.float64
        return (rng.uniform(-1, 1, n).astype(comp) + 1j * rng.uniform(-1, 1, n).astype(comp)).astype(dt)
    return rng.uniform(-1.0, 1.0, n).astype(dt)


def unique_sorted_columns(rng, lengths: np.ndarray, n_cols: int) -> np.ndarray:
    """Per-row sorted, duplicate-free uniform columns; row r gets lengths[r]
    (which must be <= n_cols).  Duplicates from sampling with replacement are
    redrawn for the affected rows only, until none remain."""
    lengths = np.asarray(lengths, dtype=np.int64)
    n = lengths.size
    nnz = int(lengths.sum())
    if nnz == 0:
        return np.zeros(0, dtype=np.int64)
    rows = np.repeat(np.arange(n, dtype=np.int64), lengths)
    keys = rows * n_cols + rng.integers(0, n_cols, nnz, dtype=np.int64)
    # rows denser than 1/8 draw without replacement directly (redrawing
    # duplicates would take ~n_cols rounds for a nearly full row)
    dense = np.nonzero(lengths * 8 > n_cols)[0]
    if dense.size:
        starts = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(lengths, out=starts[1:])
        for r in dense:
            keys[starts[r]:starts[r + 1]] = r * n_cols + rng.choice(n_cols, int(lengths[r]),
                                                                    replace=False)
    keys.sort()
    while True:
        dup = np.zeros(nnz, dtype=bool)
        dup[1:] = keys[1:] == keys[:-1]
        if not dup.any():
            break
        bad = np.zeros(n, dtype=bool)
        bad[keys[dup] // n_cols] = True
        sel = bad[rows]
        sub = keys[sel]
        sub_dup = dup[sel]
        sub_rows = sub // n_cols
        sub[sub_dup] = sub_rows[sub_dup] * n_cols + rng.integers(0, n_cols, int(sub_dup.sum()),
                                                                dtype=np.int64)
        sub.sort()
        keys[sel] = sub
    return keys - rows * n_cols


def _assemble(lengths, cols, vals):
    indptr = np.zeros(lengths.size + 1, dtype=np.int64)
    np.cumsum(lengths, out=indptr[1:])
    return indptr, cols, vals


def poisson_csr(n, mean, seed, lo=1, dtype=np.float32, n_cols=None, index_dtype=np.int32):
    rng = _rng(seed)
    n_cols = n if n_cols is None else n_cols
    lengths = np.clip(rng.poisson(mean, n), lo, n_cols).astype(np.int64)
    cols = unique_sorted_columns(rng, lengths, n_cols).astype(index_dtype)
    return _assemble(lengths, cols, _values(rng, cols.size, dtype))


def powerlaw_lengths(rng, n, alpha=2.0, min_len=1, max_len=100_000):
    """ceil of a continuous power law ~ l^-alpha on [min_len, max_len]."""
    u = rng.random(n)
    if abs(alpha - 1.0) < 1e-12:
        l = min_len * (max_len / min_len) ** u
    else:
        a = 1.0 - alpha
        l = (min_len ** a + u * (max_len ** a - min_len ** a)) ** (1.0 / a)
    return np.clip(np.ceil(l), min_len, max_len).astype(np.int64)


def powerlaw_csr(n, seed, alpha=2.0, max_len=100_000, dtype=np.float32, index_dtype=np.int32):
    rng = _rng(seed)
    lengths = np.minimum(powerlaw_lengths(rng, n, alpha, 1, max_len), n)
    cols = unique_sorted_columns(rng, lengths, n).astype(index_dtype)
    return _assemble(lengths, cols, _values(rng, cols.size, dtype))


def banded_random_csr(n, seed, half_band=3, n_random=8, dtype=np.float32, index_dtype=np.int32):
    """Offsets -h..h (clipped) plus n_random distinct uniform off-band columns."""
    rng = _rng(seed)
    r = np.arange(n, dtype=np.int64)
    band = r[:, None] + np.arange(-half_band, half_band + 1, dtype=np.int64)[None, :]
    band_ok = (band >= 0) & (band < n)
    # off-band draws: uniform over the n - (band width) columns outside the band
    extra = np.empty((n, n_random), dtype=np.int64)
    todo = np.arange(n)
    while todo.size:
        k = todo.size
        draw = rng.integers(0, n, (k, n_random), dtype=np.int64)
        rr = r[todo][:, None]
        in_band = np.abs(draw - rr) <= half_band
        srt = np.sort(draw, axis=1)
        dup = np.zeros_like(draw, dtype=bool)
        dup[:, 1:] = srt[:, 1:] == srt[:, :-1]
        bad = in_band.any(axis=1) | dup.any(axis=1)
        extra[todo[~bad]] = srt[~bad]
        todo = todo[bad]
    lengths = band_ok.sum(axis=1) + n_random
    cols = np.concatenate([band, extra], axis=1)
    keep = np.concatenate([band_ok, np.ones((n, n_random), dtype=bool)], axis=1)
    cols = np.where(keep, cols, np.iinfo(np.int64).max)
    cols.sort(axis=1)
    flat = cols[cols != np.iinfo(np.int64).max]
    return _assemble(lengths.astype(np.int64), flat.astype(index_dtype),
                     _values(rng, flat.size, dtype))


def banded_csr(n, seed, bandwidth, dtype=np.float32, index_dtype=np.int32):
    """Full band of 2*bandwidth+1 diagonals (clipped at the edges)."""
    rng = _rng(seed)
    r = np.arange(n, dtype=np.int64)
    lo = np.maximum(r - bandwidth, 0)
    hi = np.minimum(r + bandwidth, n - 1)
    lengths = hi - lo + 1
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lengths, out=indptr[1:])
    cols = np.arange(indptr[-1], dtype=np.int64) - np.repeat(indptr[:-1], lengths) + np.repeat(lo, lengths)
    return indptr, cols.astype(index_dtype), _values(rng, cols.size, dtype)


def block_diag_csr(n, seed, block=8, dtype=np.float32, index_dtype=np.int32):
    """Dense block x block blocks on the diagonal (last block clipped)."""
    rng = _rng(seed)
    r = np.arange(n, dtype=np.int64)
    lo = (r // block) * block
    hi = np.minimum(lo + block, n)
    lengths = hi - lo
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lengths, out=indptr[1:])
    cols = np.arange(indptr[-1], dtype=np.int64) - np.repeat(indptr[:-1], lengths) + np.repeat(lo, lengths)
    return indptr, cols.astype(index_dtype), _values(rng, cols.size, dtype)


def rmat_csr(n, nnz_target, seed, a=0.57, b=0.19, c=0.19, dtype=np.float32,
             index_dtype=np.int32):
    """R-MAT edges on a 2^k grid folded into n rows / cols, duplicates merged
    by keeping one entry per (row, col)."""
    rng = _rng(seed)
    k = max(1, math.ceil(math.log2(max(n, 2))))
    rows = np.zeros(nnz_target, dtype=np.int64)
    cols = np.zeros(nnz_target, dtype=np.int64)
    p = np.array([a, b, c, 1.0 - a - b - c])
    for bit in range(k):
        q = rng.choice(4, size=nnz_target, p=p)
        rows |= ((q >> 1) & 1).astype(np.int64) << bit
        cols |= (q & 1).astype(np.int64) << bit
    rows %= n
    cols %= n
    keys = np.unique(rows * n + cols)
    r, cc = keys // n, keys % n
    lengths = np.bincount(r, minlength=n).astype(np.int64)
    return _assemble(lengths, cc.astype(index_dtype), _values(rng, cc.size, dtype))


def dense_uniform(rng_or_seed, shape, dtype=np.float32):
    rng = rng_or_seed if isinstance(rng_or_seed, np.random.Generator) else _rng(rng_or_seed)
    return _values(rng, int(np.prod(shape)), dtype).reshape(shape)


# Named workloads ------------------------------------------------------------

def spmv_16k(seed=1, dtype=np.float32):
    """Square n = 16384, lengths clip(Poisson(16), 1, n)."""
    return poisson_csr(16384, 16.0, seed, lo=1, dtype=dtype)


def spmv_powerlaw_4m(seed=2, dtype=np.float32, n=4_000_000):
    """n = 4,000,000, power-law lengths (alpha 2, max 1e5)."""
    return powerlaw_csr(n, seed, alpha=2.0, max_len=100_000, dtype=dtype)


def spmm_banded_1m(seed=3, dtype=np.float32, n=1_048_576):
    """n = 1,048,576, band -3..+3 plus 8 off-band columns."""
    return banded_random_csr(n, seed, half_band=3, n_random=8, dtype=dtype)


# Device-side generators (torch, seeded philox on the GPU) for the selector's
# matrix families.  Columns are uniform with replacement (duplicates summed as
# stored) except where the family fixes them; rows are sorted by column.

SELECTOR_FAMILIES = ("poisson", "powerlaw", "banded", "blockdiag", "rmat")
# Families whose rows' columns lie within a few dozen of the diagonal
# (rowsplit.columns_local is true for them); the others draw columns at random.
LOCAL_FAMILIES = ("banded", "blockdiag")


def _torch_rows_to_csr(torch, rows_len, cols, n, vals_gen, device):
    indptr = torch.zeros(n + 1, dtype=torch.int64, device=device)
    torch.cumsum(rows_len, 0, out=indptr[1:])
    nnz = int(indptr[-1].item())
    vals = (torch.rand(nnz, generator=vals_gen, device=device) * 2 - 1).float()
    return indptr, cols.to(torch.int32), vals


def _torch_sorted_random_cols(torch, lens, n_cols, gen, device):
    n = lens.numel()
    nnz = int(lens.sum().item())
    rows = torch.repeat_interleave(torch.arange(n, device=device), lens)
    cols = torch.randint(0, n_cols, (nnz,), generator=gen, device=device)
    key = rows * n_cols + cols
    key, _ = torch.sort(key)
    return key - rows * n_cols


def family_csr_torch(family, n, seed, device, params=None):
    """One matrix of a selector family on the device; returns (indptr int64,
    indices int32, values fp32, params dict).  params not given are drawn from
    the seed: poisson mean ~ U[4, 32]; power-law alpha ~ U[1.5, 3] (max length
    min(n, 1e5)); band width ~ U{3..63} (offsets -w//2..w//2); block 8;
    R-MAT (a, b, c) = (0.57, 0.19, 0.19) with n * U[4, 24] edges."""
    import torch

    rng = np.random.Generator(np.random.PCG64(seed))
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    p = dict(params or {})
    if family == "poisson":
        mean = p.setdefault("mean", float(rng.uniform(4, 32)))
        lens = torch.poisson(torch.full((n,), mean, device=device), generator=gen).to(torch.int64)
        lens = torch.clamp(lens, 0, n)
        cols = _torch_sorted_random_cols(torch, lens, n, gen, device)
    elif family == "powerlaw":
        alpha = p.setdefault("alpha", float(rng.uniform(1.5, 3.0)))
        max_len = min(n, 100_000)
        u = torch.rand(n, generator=gen, device=device, dtype=torch.float64)
        a = 1.0 - alpha
        l = (1.0 + u * (max_len ** a - 1.0)) ** (1.0 / a)
        lens = torch.clamp(torch.ceil(l), 1, max_len).to(torch.int64)
        cols = _torch_sorted_random_cols(torch, lens, n, gen, device)
    elif family == "banded":
        w = p.setdefault("width", int(rng.integers(3, 64)))
        h = w // 2
        r = torch.arange(n, device=device)
        lo = torch.clamp(r - h, min=0)
        hi = torch.clamp(r + h, max=n - 1)
        lens = hi - lo + 1
        start = torch.zeros(n + 1, dtype=torch.int64, device=device)
        torch.cumsum(lens, 0, out=start[1:])
        rows = torch.repeat_interleave(r, lens)
        cols = torch.arange(int(start[-1].item()), device=device) - start[rows] + lo[rows]
    elif family == "blockdiag":
        b = p.setdefault("block", 8)
        r = torch.arange(n, device=device)
        lo = (r // b) * b
        lens = torch.clamp(lo + b, max=n) - lo
        start = torch.zeros(n + 1, dtype=torch.int64, device=device)
        torch.cumsum(lens, 0, out=start[1:])
        rows = torch.repeat_interleave(r, lens)
        cols = torch.arange(int(start[-1].item()), device=device) - start[rows] + lo[rows]
    elif family == "rmat":
        deg = p.setdefault("degree", float(rng.uniform(4, 24)))
        m = int(n * deg)
        k = max(1, int(np.ceil(np.log2(max(n, 2)))))
        probs = torch.tensor([0.57, 0.19, 0.19, 0.05], device=device)
        rr = torch.zeros(m, dtype=torch.int64, device=device)
        cc = torch.zeros(m, dtype=torch.int64, device=device)
        for bit in range(k):
            q = torch.multinomial(probs, m, replacement=True, generator=gen)
            rr |= ((q >> 1) & 1) << bit
            cc |= (q & 1) << bit
        key = torch.unique((rr % n) * n + (cc % n))
        rows, cols = key // n, key % n
        lens = torch.bincount(rows, minlength=n)
    else:
        raise ValueError(f"unknown family {family}")
    indptr, cols, vals = _torch_rows_to_csr(torch, lens, cols, n, gen, device)
    return indptr, cols, vals, p


SELECTOR_LARGE_SEED0 = 9000  # seeds from here on draw the large-matrix supplement


def selector_matrix_spec(seed):
    """(family, n) of selector matrix `seed`: family by seed mod 5, n
    log-uniform in [1e4, 2e6] (seeds below SELECTOR_LARGE_SEED0) or in
    [2e6, 5e6] (the large-matrix supplement)."""
    rng = np.random.Generator(np.random.PCG64(10_000_000 + seed))
    family = SELECTOR_FAMILIES[seed % len(SELECTOR_FAMILIES)]
    lo, hi = (2e6, 5e6) if seed >= SELECTOR_LARGE_SEED0 else (1e4, 2e6)
    n = int(round(float(np.exp(rng.uniform(np.log(lo), np.log(hi))))))
    return family, n
