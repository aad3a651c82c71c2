"""fs_csr_analyze statistics, exactly, against numpy counts of the same
arrays: row-pointer ends, decreasing rows, column range, row-length extremes,
sum of squares, empty rows, the length histogram and the entries of long
rows.  Column arrays start at every 4-byte offset of a 16-byte line (the
pass reads the aligned body four columns at a time) and run from empty to
sizes past the eight-loads-in-flight body."""

import numpy as np
import pytest
import torch

from flagsparse_public_amd import _lib

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)
EDGES = (4, 16, 64, 256, 4096)


def _expected(ip, cols):
    n = ip.size - 1
    ln = np.diff(ip.astype(np.int64))
    st = np.zeros(_lib.FS_STATS_LEN, np.int64)
    st[0] = ip[0]
    st[1] = ip[-1]
    st[2] = int((ln < 0).sum())
    st[3] = int(cols.min()) if cols.size else np.iinfo(np.int64).max
    st[4] = int(cols.max()) if cols.size else -1
    st[5] = max(int(ln.max()), 0) if n else 0
    st[6] = int(ln.min()) if n else 0
    st[7] = int((ln * ln).sum())
    st[8] = int((ln == 0).sum())
    pos = ln[ln > 0]
    lo = 0
    for b, hi in enumerate(EDGES):
        st[9 + b] = int(((pos > lo) & (pos <= hi)).sum())
        lo = hi
    st[14] = int((pos > EDGES[-1]).sum())
    st[15] = int(pos[pos > 256].sum())
    return st


def _run(ip, cols, offset, n_cols, idx_dtype, ptr_dtype):
    c_full = torch.from_numpy(np.concatenate([np.zeros(offset, cols.dtype), cols])) \
        .to(idx_dtype).to(DEV)
    c = c_full[offset:]
    p = torch.from_numpy(ip).to(ptr_dtype).to(DEV)
    vals = torch.zeros(cols.size, dtype=torch.float32, device=DEV)
    csr = _lib.make_csr(vals, c, p, ip.size - 1, n_cols)
    return _lib.analyze(csr, DEV)


@pytest.mark.parametrize("nnz", [0, 1, 3, 5, 17, 1000, 33_000, 2_000_003])
@pytest.mark.parametrize("offset", [0, 1, 2, 3])
def test_statistics_exact(nnz, offset):
    rng = np.random.default_rng(nnz * 7 + offset)
    n = max(1, nnz // 9)
    cuts = np.sort(rng.integers(0, nnz + 1, n - 1))
    ip = np.concatenate([[0], cuts, [nnz]]).astype(np.int64)
    if n > 3:  # one long row, for the upper buckets
        ip[2:] = np.maximum(ip[2:], min(nnz, 5000))
    cols = rng.integers(0, 1 << 20, nnz).astype(np.int32)
    if nnz > 2:
        cols[rng.integers(0, nnz)] = -7        # out of range low
        cols[rng.integers(0, nnz)] = 1 << 30   # out of range high
    got = _run(ip, cols, offset, 1 << 20, torch.int32, torch.int32)
    assert got == _expected(ip, cols).tolist()


@pytest.mark.parametrize("idx_dtype,ptr_dtype", [(torch.int64, torch.int32),
                                                 (torch.int32, torch.int64),
                                                 (torch.int64, torch.int64)])
def test_statistics_wide_types_and_bad_pointers(idx_dtype, ptr_dtype):
    rng = np.random.default_rng(5)
    nnz, n = 300_001, 20_000
    ip = np.concatenate([[0], np.sort(rng.integers(0, nnz + 1, n - 1)), [nnz]]).astype(np.int64)
    ip[100], ip[101] = ip[101] + 3, ip[100]  # decreasing rows
    ip[0] = 2                                # does not start at zero
    cols = rng.integers(0, 50_000, nnz).astype(np.int64)
    if idx_dtype == torch.int64:
        cols[17] = (1 << 40) + 3
    got = _run(ip, cols, 1, 50_000, idx_dtype, ptr_dtype)
    assert got == _expected(ip, cols).tolist()
