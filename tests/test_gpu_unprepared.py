"""Unprepared SpMV calls (flagsparse_spmv_csr(data, indices, indptr, x,
shape)): a repeated call on the same tensors reuses the prepared schedule and
layouts after a fresh validation pass; any change to the tensors (in-place
writes through torch, a new tensor, another shape or option, a structure
change that the validation statistics see) prepares again; results always
match the oracle on the current contents and invalid contents still raise.
"""

import importlib

import numpy as np
import pytest
import torch

from flagsparse_public_amd import flagsparse_spmv_csr, synthetic
from tests._util import host64, row_rel_err, to_dev

pytestmark = pytest.mark.gpu

spmv_mod = importlib.import_module("flagsparse_public_amd.sparse_operations.spmv_csr")


def _matrix(n=20_000, seed=31):
    ip, cols, vals = synthetic.powerlaw_csr(n, seed, alpha=2.0, max_len=3000)
    return ip, cols, vals


def _check(orc, y, ip, cols, vals, x):
    y64 = orc.spmv(ip, cols, vals, x)
    cond = orc.spmv_abs(ip, cols, vals, x)
    assert row_rel_err(host64(y), y64, cond).max() <= 1e-5


def _entry(data):
    """The newest cache entry for these values (None if none)."""
    for k, refs, prepared in reversed(spmv_mod._UNPREPARED._entries):
        if refs[0]() is data:
            return prepared
    return None


def test_repeated_calls_reuse_the_prepared_matrix(orc):
    spmv_mod._UNPREPARED.clear()
    ip, cols, vals = _matrix()
    n = ip.size - 1
    x = synthetic.dense_uniform(1, n)
    d, c, p, dx = to_dev(vals), to_dev(cols), to_dev(ip), to_dev(x)
    y1 = flagsparse_spmv_csr(d, c, p, dx, (n, n), use_opt=True)
    first = _entry(d)
    assert first is not None
    y2 = flagsparse_spmv_csr(d, c, p, dx, (n, n), use_opt=True)
    assert _entry(d) is first and torch.equal(y1, y2)
    _check(orc, y2, ip, cols, vals, x)
    # a changed option is another entry
    flagsparse_spmv_csr(d, c, p, dx, (n, n), use_opt=True, block_nnz=512)
    assert len(spmv_mod._UNPREPARED._entries) == 2


def test_in_place_value_write_prepares_again(orc):
    spmv_mod._UNPREPARED.clear()
    ip, cols, vals = _matrix(seed=32)
    n = ip.size - 1
    x = synthetic.dense_uniform(2, n)
    d, c, p, dx = to_dev(vals), to_dev(cols), to_dev(ip), to_dev(x)
    flagsparse_spmv_csr(d, c, p, dx, (n, n))
    first = _entry(d)
    d.mul_(-3.0)  # bumps d's version counter
    y = flagsparse_spmv_csr(d, c, p, dx, (n, n), use_opt=True)
    assert _entry(d) is not first
    assert sum(refs[0]() is d for _, refs, _ in spmv_mod._UNPREPARED._entries) == 1
    _check(orc, y, ip, cols, (vals * np.float32(-3.0)).astype(np.float32), x)


def test_structure_changes_are_seen(orc):
    spmv_mod._UNPREPARED.clear()
    ip, cols, vals = _matrix(seed=33)
    n = ip.size - 1
    x = synthetic.dense_uniform(3, n)
    d, c, p, dx = to_dev(vals), to_dev(cols), to_dev(ip), to_dev(x)
    flagsparse_spmv_csr(d, c, p, dx, (n, n), use_opt=True)
    first = _entry(d)
    # a write that bypasses the version counter but moves a row boundary:
    # the validation statistics (row-length histogram) differ -> prepared again
    ip2 = ip.copy()
    k = int(np.argmax(np.diff(ip) > 2))
    ip2[k + 1] -= 1
    p.data.copy_(to_dev(ip2))
    y = flagsparse_spmv_csr(d, c, p, dx, (n, n), use_opt=True)
    assert _entry(d) is not first
    _check(orc, y, ip2, cols, vals, x)
    # an out-of-range column written in place still raises
    c.data[0] = n + 5
    with pytest.raises(IndexError):
        flagsparse_spmv_csr(d, c, p, dx, (n, n), use_opt=True)


def test_transposed_unprepared_calls(orc):
    spmv_mod._UNPREPARED.clear()
    ip, cols, vals = _matrix(n=8000, seed=34)
    n = ip.size - 1
    x = synthetic.dense_uniform(4, n)
    d, c, p, dx = to_dev(vals), to_dev(cols), to_dev(ip), to_dev(x)
    y1 = flagsparse_spmv_csr(d, c, p, dx, (n, n), op="trans", use_opt=True)
    y2 = flagsparse_spmv_csr(d, c, p, dx, (n, n), op="trans", use_opt=True)
    assert torch.equal(y1, y2)
    t_ip, t_cols, perm = orc.csr_transpose(ip, cols, (n, n))
    _check(orc, y2, t_ip, t_cols[:int(t_ip[-1])], vals[perm[:int(t_ip[-1])]], x)
