"""The oracle's SDDMM pinned by a different route: the dense product X Y^T
(numpy matmul) sampled at the pattern, plus a hand known answer."""

import numpy as np


def test_sddmm_oracle_equals_sampled_dense_product(orc):
    rng = np.random.default_rng(3)
    n, m, k = 40, 30, 17
    mask = rng.random((n, m)) < 0.2
    rows, cols = np.nonzero(mask)
    ip = np.zeros(n + 1, np.int64)
    np.add.at(ip, rows + 1, 1)
    ip = np.cumsum(ip)
    x = rng.uniform(-1, 1, (n, k))
    y = rng.uniform(-1, 1, (m, k))
    data = rng.uniform(-1, 1, rows.size)
    dense = x @ y.T
    out, cond = orc.sddmm(ip, cols, x, y, data, 0.75, -2.0)
    want = 0.75 * dense[rows, cols] - 2.0 * data
    assert np.allclose(out, want, rtol=1e-13, atol=1e-13)
    assert np.all(cond + 1e-15 >= np.abs(out))


def test_sddmm_oracle_known_answer(orc):
    ip = np.array([0, 2, 3])
    ix = np.array([0, 2, 1])
    x = np.array([[1.0, 2.0], [3.0, 4.0]])
    y = np.array([[1.0, 0.0], [0.0, 1.0], [2.0, 2.0]])
    out, _ = orc.sddmm(ip, ix, x, y, np.ones(3), 2.0, 0.5)
    assert out.tolist() == [2.5, 12.5, 8.5]
    # rows of x against rows of y: swapping the operands' roles gives another answer
    swapped = [2.0 * (y[0] @ x[0]) + 0.5, 2.0 * (y[0] @ np.array([2.0, 2.0])) + 0.5]
    assert swapped[1] != out[1]
