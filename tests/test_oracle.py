"""Pins the CPU oracle (CPU only).

Every oracle function is checked against something that is not its own
formula: SciPy's CSR routines (bit-for-bit, same summation order), the
exactly rounded golden fixtures, hand-derived known answers from the
reference's bucket tables, finite differences and closed forms.
"""

import math

import numpy as np
import pytest
import scipy.sparse as sp

GOLDEN = None


def golden():
    global GOLDEN
    if GOLDEN is None:
        import os
        GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "csr_golden.npz"))
    return GOLDEN


def golden_case(name):
    g = golden()
    pre = name + "__"
    return {k[len(pre):]: g[k] for k in g.files if k.startswith(pre)}


GOLDEN_NAMES = ["ka2x2", "longrow", "unsorted_dup", "empty_rows", "all_empty", "cancel", "complex"]


def rand_csr(rng, n, m, density, dup=False, unsorted=False, dtype=np.float64):
    lens = rng.poisson(density * m, n)
    lens[rng.random(n) < 0.1] = 0
    ip = np.zeros(n + 1, np.int64)
    np.cumsum(lens, out=ip[1:])
    cols = rng.integers(0, m, ip[-1]) if dup else None
    if not dup:
        cols = np.concatenate([np.sort(rng.choice(m, min(l, m), replace=False)) for l in lens]) \
            if ip[-1] else np.zeros(0, np.int64)
        lens = np.minimum(lens, m)
        np.cumsum(lens, out=ip[1:])
    if unsorted:
        for r in range(n):
            seg = cols[ip[r]:ip[r + 1]]
            rng.shuffle(seg)
    vals = rng.uniform(-1, 1, ip[-1]).astype(dtype)
    return ip, cols.astype(np.int64), vals


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("dup,unsorted", [(False, False), (True, True)])
def test_spmv_bitexact_vs_scipy(orc, seed, dup, unsorted):
    rng = np.random.default_rng(seed)
    ip, cols, vals = rand_csr(rng, 300, 200, 0.05, dup=dup, unsorted=unsorted)
    x = rng.uniform(-1, 1, 200)
    A = sp.csr_matrix((vals, cols, ip), shape=(300, 200))
    ref = A @ x  # SciPy's csr_matvec: sequential sum in stored order
    got = orc.spmv(ip, cols, vals, x)
    assert np.array_equal(got, ref)


def test_spmv_complex_matches_scipy(orc):
    rng = np.random.default_rng(5)
    ip, cols, vals = rand_csr(rng, 120, 90, 0.1, dup=True)
    cv = vals + 1j * rng.uniform(-1, 1, vals.size)
    x = rng.uniform(-1, 1, 90) + 1j * rng.uniform(-1, 1, 90)
    ref = sp.csr_matrix((cv, cols, ip), shape=(120, 90)) @ x
    got = orc.spmv(ip, cols, cv, x)
    # same products and order; allow only fused-vs-unfused rounding differences
    assert np.max(np.abs(got - ref)) <= 1e-15 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("nd", [1, 5, 32])
def test_spmm_bitexact_vs_scipy(orc, nd):
    rng = np.random.default_rng(7)
    ip, cols, vals = rand_csr(rng, 150, 100, 0.08, dup=True, unsorted=True)
    B = rng.uniform(-1, 1, (100, nd))
    ref = sp.csr_matrix((vals, cols, ip), shape=(150, 100)) @ B
    got = orc.spmm(ip, cols, vals, B)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("nd", [1, 5, 32])
def test_spmm_complex_matches_scipy(orc, nd):
    """Complex SpMM against SciPy's complex csr_matvecs (same products, same
    order): only fused-vs-unfused rounding may differ."""
    rng = np.random.default_rng(17 + nd)
    ip, cols, vals = rand_csr(rng, 130, 70, 0.1, dup=True, unsorted=True)
    cv = vals + 1j * rng.uniform(-1, 1, vals.size)
    B = rng.uniform(-1, 1, (70, nd)) + 1j * rng.uniform(-1, 1, (70, nd))
    ref = sp.csr_matrix((cv, cols, ip), shape=(130, 70)) @ B
    got = orc.spmm(ip, cols, cv, B)
    assert got.shape == ref.shape and got.dtype == np.complex128
    assert np.max(np.abs(got - ref)) <= 1e-15 * max(1.0, np.max(np.abs(ref)))
    # a column of B is the complex SpMV with that column
    for k in (0, nd - 1):
        yk = orc.spmv(ip, cols, cv, B[:, k])
        assert np.max(np.abs(got[:, k] - yk)) <= 1e-15 * max(1.0, np.max(np.abs(yk)))


def test_spmm_complex_known_answer(orc):
    """[[1+2j, 3-1j]] @ [[2-1j, 1], [1+1j, 1j]] = [[8+5j, 2+5j]], exactly: a
    dropped or sign-flipped imaginary cross term changes it."""
    ip = np.array([0, 2])
    cols = np.array([0, 1])
    cv = np.array([1 + 2j, 3 - 1j])
    B = np.array([[2 - 1j, 1 + 0j], [1 + 1j, 1j]])
    C = orc.spmm(ip, cols, cv, B)
    assert np.array_equal(C, np.array([[8 + 5j, 2 + 5j]]))
    assert np.array_equal(orc.spmv(ip, cols, cv, B[:, 0]), np.array([8 + 5j]))


@pytest.mark.parametrize("dup,unsorted", [(False, False), (True, True)])
def test_condition_sums_bitexact_vs_scipy(orc, dup, unsorted):
    """sum_j |a_ij x_j| and sum_j |a_ij B_jk| are |A| |x| and |A| |B|: for
    real fp64 values |a x| = |a| |x| exactly, and SciPy sums in the same
    order, so the results are bit-identical."""
    rng = np.random.default_rng(23)
    ip, cols, vals = rand_csr(rng, 240, 150, 0.06, dup=dup, unsorted=unsorted)
    x = rng.uniform(-1, 1, 150)
    B = rng.uniform(-1, 1, (150, 9))
    # built from |data| directly: SciPy's abs() first merges repeated entries
    absA = sp.csr_matrix((np.abs(vals), cols, ip), shape=(240, 150))
    assert np.array_equal(orc.spmv_abs(ip, cols, vals, x), absA @ np.abs(x))
    assert np.array_equal(orc.spmm_abs(ip, cols, vals, B), absA @ np.abs(B))
    # not the magnitude of the sum: cancellation shows the difference
    assert np.all(orc.spmv_abs(ip, cols, vals, x) >= np.abs(orc.spmv(ip, cols, vals, x)))


def test_condition_sums_complex_vs_scipy(orc):
    """Complex: |a x| = |a| |x| up to the rounding of the two moduli, so the
    sums agree with SciPy's |A| |x| within a few ulps per term."""
    rng = np.random.default_rng(29)
    ip, cols, vals = rand_csr(rng, 160, 110, 0.08, dup=True)
    cv = vals + 1j * rng.uniform(-1, 1, vals.size)
    x = rng.uniform(-1, 1, 110) + 1j * rng.uniform(-1, 1, 110)
    B = rng.uniform(-1, 1, (110, 6)) + 1j * rng.uniform(-1, 1, (110, 6))
    absA = sp.csr_matrix((np.abs(cv), cols, ip), shape=(160, 110))
    ref_v = absA @ np.abs(x)
    ref_m = absA @ np.abs(B)
    got_v = orc.spmv_abs(ip, cols, cv, x)
    got_m = orc.spmm_abs(ip, cols, cv, B)
    assert got_v.dtype == np.float64 and got_m.dtype == np.float64
    assert np.all(np.abs(got_v - ref_v) <= 8 * 2.0 ** -53 * ref_v)
    assert np.all(np.abs(got_m - ref_m) <= 8 * 2.0 ** -53 * ref_m)


def test_condition_sum_known_answer(orc):
    """A = [[1, -2], [0, 0.5]], x = [3, 4]: y = [-5, 2], sums [11, 2]."""
    ip, cols, vals = np.array([0, 2, 3]), np.array([0, 1, 1]), np.array([1.0, -2.0, 0.5])
    x = np.array([3.0, 4.0])
    assert np.array_equal(orc.spmv(ip, cols, vals, x), [-5.0, 2.0])
    assert np.array_equal(orc.spmv_abs(ip, cols, vals, x), [11.0, 2.0])
    B = np.array([[3.0, -1.0], [4.0, 2.0]])
    assert np.array_equal(orc.spmm_abs(ip, cols, vals, B), [[11.0, 5.0], [2.0, 1.0]])
    cx = np.array([3j, 4.0])
    assert np.array_equal(orc.spmv_abs(ip, cols, vals.astype(complex), cx), [11.0, 2.0])


@pytest.mark.parametrize("name", GOLDEN_NAMES)
def test_oracle_against_exact_golden(orc, name):
    c = golden_case(name)
    ip, cols, data, x = c["indptr"], c["indices"], c["data"], c["x"]
    got = orc.spmv(ip, cols, data, x)
    cond = orc.spmv_abs(ip, cols, data, x)
    lens = np.diff(ip)
    # sequential fp64 summation error bound: (len - 1) u sum|a x| (products of
    # fp32 inputs are exact in fp64; fp64 products add one more u)
    bound = (np.maximum(lens, 1) + 1) * 2.0 ** -53 * cond
    assert np.all(np.abs(got - c["y_exact"]) <= bound + 1e-300)
    if "y_f32" in c:
        assert np.array_equal(got.astype(np.float32), c["y_f32"]) or name == "cancel"
    if "C_exact" in c:
        C = orc.spmm(ip, cols, data, c["B"])
        condC = orc.spmm_abs(ip, cols, data, c["B"])
        assert np.all(np.abs(C - c["C_exact"]) <= (lens[:, None] + 1) * 2.0 ** -53 * condC + 1e-300)


def test_known_answer_2x2(orc):
    c = golden_case("ka2x2")
    # [[1,2],[0,3]] @ [0.5, -1.25] = [-2.0, -3.75]
    assert np.array_equal(c["y_exact"], [-2.0, -3.75])
    assert np.array_equal(orc.spmv(c["indptr"], c["indices"], c["data"], c["x"]), [-2.0, -3.75])


def test_dropped_term_is_caught(orc):
    """A restatement that moves a row boundary by one entry (row 0 drops its
    last term, row 1 gains it) fails the golden check."""
    c = golden_case("longrow")
    bad = orc.spmv(np.array([0, 1535, 1540]), c["indices"], c["data"], c["x"])
    assert not np.any(np.isclose(bad, c["y_exact"], rtol=1e-12, atol=0))


@pytest.mark.parametrize("seed", [0, 3])
def test_transpose_matches_scipy_tocsc(orc, seed):
    rng = np.random.default_rng(seed)
    ip, cols, vals = rand_csr(rng, 80, 60, 0.1, dup=True, unsorted=True)
    A = sp.csr_matrix((vals, cols, ip), shape=(80, 60))
    # SciPy's tocsc keeps duplicates and orders entries of a column by row,
    # stably; canonical form is not forced.
    Ac = A.tocsc()
    t_ip, t_idx, perm = orc.csr_transpose(ip, cols, (80, 60))
    assert np.array_equal(t_ip, Ac.indptr)
    assert np.array_equal(t_idx, Ac.indices)
    assert np.array_equal(vals[perm], Ac.data)


def test_spmv_opt_bucket_known_answers(orc):
    # reference table spmv_csr.py:128-153: f32 bounds 64 / 512 / 4096 / inf,
    # f64 bounds 64 / 256 / 2048 / 8192 / inf; empty rows in bucket 0
    lens = [0, 1, 64, 65, 512, 513, 4096, 4097, 10**6]
    assert list(orc.spmv_opt_bucket(lens)) == [0, 0, 0, 1, 1, 2, 2, 3, 3]
    lens64 = [0, 64, 65, 256, 257, 2048, 2049, 8192, 8193]
    assert list(orc.spmv_opt_bucket(lens64, fp64=True)) == [0, 0, 1, 1, 2, 2, 3, 3, 4]


def test_spmm_bucket_known_answers(orc):
    # spmm_csr.py:395-401 bounds 32 / 128 / 512 / 2048 / inf
    assert list(orc.spmm_opt_bucket([0, 32, 33, 128, 129, 512, 513, 2048, 2049])) == \
        [0, 0, 1, 1, 2, 2, 3, 3, 4]
    # alg2 inclusive ranges [0,16] [17,64] [65,256] [257,1024] [1025,inf)
    assert list(orc.spmm_alg2_bucket([0, 16, 17, 64, 65, 256, 257, 1024, 1025])) == \
        [0, 0, 1, 1, 2, 2, 3, 3, 4]


def test_split_parts_known_answer(orc):
    # long row of 1536 entries -> 6 parts of 256; row of 300 -> 256 + 44
    ip = np.array([0, 1536, 1540, 1840])
    pr, ps, pe, rpp = orc.split_parts(ip, [0, 2], 256)
    assert list(pr) == [0] * 6 + [2, 2]
    assert list(ps) == [0, 256, 512, 768, 1024, 1280, 1540, 1796]
    assert list(pe) == [256, 512, 768, 1024, 1280, 1536, 1796, 1840]
    assert list(rpp) == [0, 6, 8]


def test_row_features_hand_case(orc):
    ip = np.array([0, 0, 3, 23, 23, 300])   # lengths 0, 3, 20, 0, 277
    f = orc.row_features(ip, 1000)
    lens = np.array([0, 3, 20, 0, 277])
    assert f[0] == pytest.approx(math.log2(6))
    assert f[1] == pytest.approx(math.log2(1001))
    assert f[2] == pytest.approx(math.log2(301))
    assert f[3] == pytest.approx(60.0)
    assert f[4] == pytest.approx(np.std(lens))
    assert f[5] == pytest.approx(np.std(lens) / 60.0)
    assert f[6] == pytest.approx(math.log2(278))
    assert f[7] == 0
    assert f[8] == pytest.approx(0.4)
    assert list(f[9:15]) == pytest.approx([0.2, 0.0, 0.2, 0.0, 0.2, 0.0])
    assert f[15] == pytest.approx(277 / 300)


def _mlp_params(rng, ind=16, hid=8, outd=5):
    return [rng.normal(0, 0.5, (ind, hid)), rng.normal(0, 0.1, hid),
            rng.normal(0, 0.5, (hid, hid)), rng.normal(0, 0.1, hid),
            rng.normal(0, 0.5, (hid, outd)), rng.normal(0, 0.1, outd)]


def _mlp_loss(orc, X, P, T, M):
    _, _, Y = orc.mlp_forward(X, *P)
    e = M * (Y - T)
    return 0.5 * np.sum(e * e) / Y.size


def test_mlp_gradients_match_finite_differences(orc):
    rng = np.random.default_rng(11)
    X = rng.normal(size=(7, 16))
    P = _mlp_params(rng)
    T = rng.normal(size=(7, 5))
    M = (rng.random((7, 5)) < 0.8).astype(np.float64)
    H1, H2, Y = orc.mlp_forward(X, *P)
    loss, grads = orc.mlp_backward(X, P[2], P[4], H1, H2, Y, T, M)
    assert loss == pytest.approx(_mlp_loss(orc, X, P, T, M), rel=1e-12)
    h = 1e-6
    for pi, (p, g) in enumerate(zip(P, grads)):
        flat = p.reshape(-1)
        for k in rng.choice(flat.size, min(6, flat.size), replace=False):
            old = flat[k]
            flat[k] = old + h
            lp = _mlp_loss(orc, X, P, T, M)
            flat[k] = old - h
            lm = _mlp_loss(orc, X, P, T, M)
            flat[k] = old
            fd = (lp - lm) / (2 * h)
            assert g.reshape(-1)[k] == pytest.approx(fd, rel=1e-5, abs=1e-9), (pi, k)


def test_adam_first_step_closed_form(orc):
    # first bias-corrected step moves each parameter by lr * g / (|g| + eps)
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    m = np.zeros(3)
    v = np.zeros(3)
    orc.adam_step(p, g, m, v, 0.01, 0.9, 0.999, 1e-8, 1)
    expect = np.array([1.0, -2.0, 0.5]) - 0.01 * g / (np.abs(g) + 1e-8)
    assert p == pytest.approx(expect, rel=1e-12)


def test_bf16_split_matches_torch_rounding(orc):
    import torch
    rng = np.random.default_rng(3)
    v = np.concatenate([rng.uniform(-1, 1, 5000), rng.normal(0, 1e4, 500),
                        [0.0, -0.0, 1.0, 3.0e38, 1e-40]]).astype(np.float32)
    hi, lo = orc.bf16_split(v)
    t = torch.from_numpy(v).to(torch.bfloat16)
    assert np.array_equal(hi, t.view(torch.int16).numpy().view(np.uint16))
    r = v - torch.from_numpy(hi.view(np.int16)).view(torch.bfloat16).float().numpy()
    assert np.array_equal(lo, torch.from_numpy(r).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16))
    joined = orc.bf16_join(hi, lo)
    finite = np.abs(v) < 1e38
    rel = np.abs(joined[finite] - v[finite]) / np.maximum(np.abs(v[finite]), 1e-30)
    assert np.all(rel <= 2.0 ** -16)


def test_allclose_ratio_definition(orc):
    # max(|d| / (atol + rtol |ref|)), tests/test_spmv.py:179-184
    assert orc.allclose_ratio([1.0, 2.0], [1.0, 2.5], 0.1, 0.1) == pytest.approx(0.5 / 0.35)
