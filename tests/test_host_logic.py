"""Host-side logic of the operator modules (CPU only): op codes, validation
order and messages, bucket / split / transpose metadata against the oracle,
the int64 -> int32 fallback contract, schedules and the synthetic workloads."""

import importlib

import numpy as np
import pytest
import torch

spmv_mod = importlib.import_module("flagsparse_public_amd.sparse_operations.spmv_csr")
spmm_mod = importlib.import_module("flagsparse_public_amd.sparse_operations.spmm_csr")
alg2_mod = importlib.import_module("flagsparse_public_amd.sparse_operations.spmm_csr_opt_alg2")
from flagsparse_public_amd import rowsplit, synthetic  # noqa: E402


@pytest.mark.parametrize("op,transpose,code", [
    (None, False, 0), (None, True, 1), ("non", False, 0), ("TRANS", False, 1), (" conj ", False, 2),
    (0, False, 0), (1, False, 1), (2, False, 2), (1.0, False, 1)])
def test_normalize_op(op, transpose, code):
    assert spmv_mod._normalize_spmv_op(op, transpose=transpose) == code


@pytest.mark.parametrize("bad", ["herm", "1", 3, -1, object()])
def test_normalize_op_rejects(bad):
    with pytest.raises(ValueError, match="op must be one of: 0=non, 1=trans, 2=conj"):
        spmv_mod._normalize_spmv_op(bad)


def test_policy_and_conflicts():
    with pytest.raises(ValueError, match="index_fallback_policy"):
        spmv_mod._normalize_spmv_index_fallback_policy("lenient")
    d = torch.ones(1)
    with pytest.raises(ValueError, match="transpose=True conflicts with op=non"):
        spmv_mod.prepare_spmv_csr(d, d.int(), torch.tensor([0, 1]), (1, 1), transpose=True, op="non")
    with pytest.raises(ValueError, match="transpose conflicts with op"):
        spmv_mod.flagsparse_spmv_csr(d, d.int(), torch.tensor([0, 1]), d, (1, 1), transpose=False,
                                     op="trans")


def test_validation_order_on_cpu_tensors():
    """Checks that precede the device check fire first, with the reference's
    messages (spmv_csr.py:497-524); CPU tensors then fail the device check."""
    d = torch.ones(3)
    i = torch.zeros(3, dtype=torch.int32)
    p = torch.tensor([0, 3])
    with pytest.raises(TypeError, match="must all be torch.Tensor"):
        spmv_mod._prepare_spmv_csr_matrix([1.0], i, p, (1, 4))
    with pytest.raises(ValueError, match="must be 1D tensors"):
        spmv_mod._prepare_spmv_csr_matrix(d.view(3, 1), i, p, (1, 4))
    with pytest.raises(ValueError, match=r"indptr length must be n_rows\+1=3, got 2"):
        spmv_mod._prepare_spmv_csr_matrix(d, i, p, (2, 4))
    with pytest.raises(ValueError, match="same length"):
        spmv_mod._prepare_spmv_csr_matrix(d, i[:2], p, (1, 4))
    with pytest.raises(ValueError, match="must be CUDA tensors"):
        spmv_mod._prepare_spmv_csr_matrix(d, i, p, (1, 4))
    with pytest.raises(ValueError, match="shape must be a 2-tuple"):
        spmm_mod._prepare_spmm_csr_matrix(d, i, p, (1,))
    with pytest.raises(ValueError, match="non-negative"):
        spmm_mod._prepare_spmm_csr_matrix(d, i, torch.tensor([0]), (-1, 4))
    with pytest.raises(ValueError, match="B must be a 2D"):
        spmm_mod._prepare_spmm_csr_inputs(d, i, p, d, (1, 4))


def test_spmv_x_validation():
    class P:
        data = torch.ones(2)
        n_cols = 4
    with pytest.raises(TypeError, match="x must be a torch.Tensor"):
        spmv_mod._validate_spmv_x(None, P)
    with pytest.raises(ValueError, match="x must be a 1D tensor"):
        spmv_mod._validate_spmv_x(torch.ones(2, 2), P)
    with pytest.raises(ValueError, match="x must be a CUDA tensor"):
        spmv_mod._validate_spmv_x(torch.ones(4), P)


@pytest.mark.parametrize("fp64", [False, True])
def test_spmv_bucket_rows_match_oracle(orc, fp64):
    lens = torch.tensor([0, 1, 64, 65, 300, 512, 513, 2048, 2049, 4096, 4097, 9000, 7])
    buckets = spmv_mod._build_spmv_opt_buckets(lens, int(lens.max()), torch.int32, fp64=fp64)
    got = np.full(lens.numel(), -1)
    for b_i, b in enumerate(buckets):
        got[b["rows"].numpy()] = b_i
    want = orc.spmv_opt_bucket(lens.numpy(), fp64=fp64)
    # bucket list skips empty buckets: compare the partition via sorted labels
    nonempty = sorted(set(want.tolist()))
    remap = {v: k for k, v in enumerate(nonempty)}
    assert got.tolist() == [remap[v] for v in want.tolist()]
    assert all(b["rows"].dtype == torch.int32 for b in buckets)


def test_spmm_buckets_and_split_parts_match_oracle(orc):
    lens = torch.tensor([0, 32, 33, 128, 129, 512, 513, 2048, 2049, 5000, 1])
    buckets, long_rows = spmm_mod._build_spmm_opt_buckets(lens)
    want = orc.spmm_opt_bucket(lens.numpy())
    for b in buckets:
        kinds = {"batched": (0, 1), "vector": (2, 3), "split": (4,)}[b["kind"]]
        assert all(want[r] in kinds for r in b["rows"].tolist())
    assert long_rows.tolist() == [8, 9]
    indptr = torch.zeros(lens.numel() + 1, dtype=torch.int64)
    indptr[1:] = torch.cumsum(lens, 0)
    pr, ps, pe, rpp = spmm_mod._build_spmm_opt_split_metadata(indptr, long_rows, 256)
    opr, ops, ope, orpp = orc.split_parts(indptr.numpy(), long_rows.numpy(), 256)
    assert pr.tolist() == opr.tolist() and ps.tolist() == ops.tolist()
    assert pe.tolist() == ope.tolist() and rpp.tolist() == orpp.tolist()


def test_alg2_buckets_match_oracle(orc):
    lens = torch.tensor([0, 16, 17, 64, 65, 256, 257, 1024, 1025, 3])
    buckets = alg2_mod._build_spmm_opt_alg2_buckets(lens, torch.float32)
    want = orc.spmm_alg2_bucket(lens.numpy())
    labels = ["short_16", "short_64", "row_256", "seg_1024", "seg_long"]
    for b in buckets:
        assert all(labels[want[r]] == b["label"] for r in b["rows"].tolist())
    assert sum(b["rows"].numel() for b in buckets) == lens.numel()


def test_alg2_launch_resolution_on_wave64():
    props = {"warp_size": 64, "max_threads_per_block": 1024, "max_threads_per_mp": 2048,
             "shared_memory_per_block": 65536}
    b = dict(alg2_mod._SPMM_OPT_ALG2_BUCKET_SPECS_F32[4])
    b["rows"] = torch.arange(3)
    launch = alg2_mod._resolve_spmm_opt_alg2_launch(b, 64, torch.float32, props)
    assert launch["block_n"] == 64 and launch["num_warps"] == 16 and launch["num_stages"] == 1
    assert alg2_mod._select_spmm_opt_alg2_block_n(5, 128) == 8
    assert alg2_mod._select_spmm_opt_alg2_block_n(200, 32) == 32


@pytest.mark.parametrize("seed", [0, 1])
def test_transpose_matches_oracle(orc, seed):
    rng = np.random.default_rng(seed)
    n, m = 40, 30
    lens = rng.poisson(3, n)
    ip = np.zeros(n + 1, np.int64)
    np.cumsum(lens, out=ip[1:])
    cols = rng.integers(0, m, ip[-1])
    vals = rng.normal(size=ip[-1])
    d, idx, ptr, shape = spmv_mod._transpose_csr_for_spmv(
        torch.from_numpy(vals), torch.from_numpy(cols).int(), torch.from_numpy(ip), (n, m))
    t_ip, t_idx, perm = orc.csr_transpose(ip, cols, (n, m))
    assert shape == (m, n)
    assert ptr.tolist() == t_ip.tolist() and idx.tolist() == t_idx.tolist()
    assert np.array_equal(d.numpy(), vals[perm])
    assert idx.dtype == torch.int32 and ptr.dtype == torch.int64


def _direct_prepared(kernel_indices, kernel_indptr, n_cols):
    data = torch.ones(kernel_indices.numel())
    return spmv_mod.PreparedCsrSpmv(
        data=data, kernel_indices=kernel_indices, kernel_indptr=kernel_indptr,
        shape=(kernel_indptr.numel() - 1, n_cols), n_rows=kernel_indptr.numel() - 1, n_cols=n_cols,
        block_nnz=256, max_segments=1, max_row_nnz=1, opt_buckets=[], transpose=False,
        index_fallback_policy="auto")


def test_fallback_refused_when_index_exceeds_int32(monkeypatch):
    """test_spmv_csr_accuracy.py:206-232, on the host."""
    limit = spmv_mod._INDEX_LIMIT_INT32
    p = _direct_prepared(torch.tensor([limit + 1]), torch.tensor([0, 1]), limit + 2)

    def fail(_p, _x, use_opt=False):
        raise RuntimeError("forced native int64 failure")

    monkeypatch.setattr(spmv_mod, "_run_spmv_prepared", fail)
    with pytest.raises(RuntimeError, match="int32 fallback is unsafe"):
        spmv_mod._run_spmv_prepared_with_fallback(p, torch.empty(0), use_opt=False)


def test_fallback_retries_with_int32(monkeypatch):
    p = _direct_prepared(torch.tensor([0, 2, 1]), torch.tensor([0, 2, 3]), 3)
    seen = []

    def launcher(prepared, x):
        seen.append(prepared.kernel_indices.dtype)
        if prepared.kernel_indices.dtype == torch.int64:
            raise RuntimeError("forced int64 launch failure")
        return "ok"

    monkeypatch.setattr(spmv_mod, "_triton_spmv_csr_impl_prepared", launcher)
    assert spmv_mod._run_spmv_prepared_with_fallback(p, None) == "ok"
    assert seen == [torch.int64, torch.int32]
    p.index_fallback_policy = "strict"
    seen.clear()
    with pytest.raises(RuntimeError, match="forced int64 launch failure"):
        spmv_mod._run_spmv_prepared_with_fallback(p, None)


def test_fallback_blocker_messages():
    limit = spmv_mod._INDEX_LIMIT_INT32
    p = _direct_prepared(torch.tensor([0]), torch.tensor([0, limit + 5], dtype=torch.int64), 2)
    assert "nnz offset" in spmv_mod._spmv_int32_fallback_blocker(p)
    p = _direct_prepared(torch.tensor([-1]), torch.tensor([0, 1]), 2)
    assert "column index range" in spmv_mod._spmv_int32_fallback_blocker(p)


def test_alg1_launch_config():
    cfg = spmm_mod._resolve_spmm_alg1_launch_config(96, 1540)
    assert (cfg["block_n"], cfg["block_nnz"], cfg["required_segments"]) == (128, 32, 49)
    assert spmm_mod._select_spmm_alg1_warp_and_factor(5) == (8, 1)
    with pytest.raises(ValueError, match="explicit max_segments"):
        spmm_mod._resolve_spmm_alg1_launch_config(32, 1540, block_nnz=128, max_segments=4)
    assert spmm_mod._select_spmm_opt_block_n(33) == 64


def test_error_metrics():
    ref = torch.tensor([1.0, -2.0, 4.0], dtype=torch.float64)
    cand = torch.tensor([1.0, -2.5, 4.0], dtype=torch.float64)
    m = spmm_mod._spmm_validation_metrics(cand, ref)
    assert m["max_abs_error"] == 0.5 and m["max_relative_error"] == pytest.approx(0.125)
    assert m["sum_relative_error"] == pytest.approx(0.5 / 7.0)
    assert m["matches_threshold"] is False
    assert spmm_mod._spmm_opt_reference_error(cand, ref, torch.float64) == pytest.approx(
        0.5 / (1e-12 + 1e-10 * 2.0))


def test_default_schedules():
    s = rowsplit.rowgroup_spmv_schedule(4_000_000, 48_000_000)
    assert s.group_lanes == 8 and s.short_max == 128 and s.chunk_nnz % 64 == 0
    assert rowsplit.rowgroup_spmv_schedule(10, 10).group_lanes == 1
    assert rowsplit.rowgroup_spmv_schedule(10, 100000).group_lanes == 64
    d = rowsplit.default_spmv_schedule(4_000_000, 48_000_000)
    assert d == rowsplit.rowgroup_spmv_schedule(4_000_000, 48_000_000)
    # x of 16 MB (fp32): entry tiles over slices of 2 MB of x; 4 MB x: row groups
    d = rowsplit.default_spmv_schedule(4_000_000, 48_000_000, torch.float32, 4_000_000)
    assert d.x_slice_cols == 512 * 1024 and d.tile_mode == 1 and d.tile_nnz == 2048
    assert rowsplit.default_spmv_schedule(10**6, 10, torch.float32, 10**6).x_slice_cols == 0
    assert rowsplit.default_spmv_schedule(10, 10, torch.float64, 1_100_000).x_slice_cols == 256 * 1024
    # with the matrix's arrays the rule samples column locality (3M columns: 12 MB of fp32 x)
    n = 3_000_000
    r = torch.arange(n, dtype=torch.int64)
    band_ptr = torch.arange(0, 3 * n + 1, 3, dtype=torch.int64)
    band_cols = torch.clamp(r.repeat_interleave(3) + torch.tensor([-1, 0, 1]).repeat(n), 0, n - 1)
    rand_ptr = torch.arange(n + 1, dtype=torch.int64)
    rand_cols = torch.randint(0, n, (n,), generator=torch.Generator().manual_seed(3))
    for vt in (torch.float32, torch.float64):
        d = rowsplit.default_spmv_schedule(n, 3 * n, vt, n, band_ptr, band_cols.to(torch.int32))
        assert d == rowsplit.Schedule(tile_nnz=2048, tile_mode=1)
        d = rowsplit.default_spmv_schedule(n, n, vt, n, rand_ptr, rand_cols)
        assert d == rowsplit.Schedule(tile_nnz=4096, tile_mode=1, x_window=-1)
    # values without x-window kernels keep the column-sliced layout for random columns
    d = rowsplit.default_spmv_schedule(n, n, torch.complex64, n, rand_ptr, rand_cols)
    assert d.x_slice_cols == 256 * 1024 and d.x_window == 0
    # so do matrices beyond the x-window layout's 2^31 - 1 entries
    d = rowsplit.default_spmv_schedule(n, 2**31, torch.float32, n, rand_ptr, rand_cols)
    assert d.x_slice_cols == 512 * 1024 and d.x_window == 0
    r = rowsplit.row_tile_spmv_schedule(10, 10, torch.complex64)
    assert r.tile_nnz == 1024 and r.short_max <= r.tile_nnz and r.chunk_nnz % 64 == 0


def test_far_column_fraction_brute_force():
    """Against a per-row loop over the same sampled rows: empty rows are
    skipped, the diagonal of a rectangular matrix is row * n_cols // n_rows."""
    rng = np.random.default_rng(11)
    n_rows, n_cols = 50_000, 170_000
    lens = rng.integers(0, 4, n_rows)
    ptr = np.zeros(n_rows + 1, np.int64)
    np.cumsum(lens, out=ptr[1:])
    cols = rng.integers(0, n_cols, int(ptr[-1]))
    far_cols = 20_000
    got = rowsplit.far_column_fraction(torch.from_numpy(ptr), torch.from_numpy(cols), n_rows,
                                       n_cols, far_cols)
    step = max(1, n_rows // rowsplit.LOCALITY_SAMPLE_ROWS)
    far = tot = 0
    for i in range(0, n_rows, step):
        if ptr[i + 1] > ptr[i]:
            tot += 1
            far += abs(int(cols[ptr[i]]) - i * n_cols // n_rows) > far_cols
    assert got == far / tot
    assert rowsplit.far_column_fraction(torch.zeros(3, dtype=torch.int64),
                                        torch.zeros(0, dtype=torch.int32), 2, 2, 1) is None


@pytest.mark.parametrize("family", synthetic.SELECTOR_FAMILIES)
def test_columns_local_by_family(family):
    """The locality sample separates the selector's families: banded and
    block-diagonal rows start at their diagonal, the random families'
    (uniform, power-law, R-MAT) first columns are spread over the row range.
    Checked at n = 200,000 with a 4,096-column band; at the large-matrix
    supplement's sizes (2M-5M rows, 32,768 columns) the sampled fractions
    are 0.000 for the local families and 0.97-0.99 for the others."""
    n = 200_000
    ip, cols, _, _ = synthetic.family_csr_torch(family, n, 7100, "cpu")
    far = rowsplit.far_column_fraction(ip, cols, n, n, 4096)
    if family in synthetic.LOCAL_FAMILIES:
        assert far == 0.0
    else:
        assert far > 0.9


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_tile_starts_invariants(seed):
    rng = np.random.default_rng(seed)
    lens = rng.poisson(3, 5000)
    lens[rng.random(5000) < 0.01] = rng.integers(100, 20000, int((rng.random(5000) < 0.01).sum()) or 1)[0]
    lens[rng.random(5000) < 0.2] = 0
    t = torch.from_numpy(lens)
    for short_max, tile in ((64, 256), (1024, 1024), (2048, 2048)):
        ts = rowsplit.tile_starts(t, short_max, tile).tolist()
        assert ts[0] == 0 and ts[-1] == lens.size and ts == sorted(set(ts))
        for a, b in zip(ts[:-1], ts[1:]):
            seg = lens[a:b]
            assert b - a <= rowsplit.TILE_ROWS
            if (seg > short_max).any():
                assert b - a == 1  # a long row is a tile of its own
            else:
                assert seg.sum() < tile + short_max


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("T", [1024, 2048, 4096])
def test_entry_tile_ownership(seed, T):
    """Entry tiles, brute force over rows and entries: the entry ranges
    partition [0, nnz) in pieces of at most T, the owned-row ranges partition
    the rows in pieces of at most 1024, every owned row starts inside its
    tile's entry range (an empty row may sit on the end), every owned row
    but the last ends inside it, and the cuts fall every T entries plus at
    row starts inside row-heavy tiles."""
    rng = np.random.default_rng(seed)
    lens = rng.poisson(5, 6000)
    lens[rng.random(6000) < 0.3] = 0
    lens[rng.integers(0, 6000, 4)] = [9000, 4096, 2 * T, 1]
    lens[1000:4000:2] = 0  # a long stretch of mostly empty rows
    lens = np.concatenate([lens, np.zeros(20000, np.int64)])  # trailing empty rows
    ip = np.zeros(lens.size + 1, np.int64)
    np.cumsum(lens, out=ip[1:])
    nnz = int(ip[-1])
    n = lens.size
    sch = rowsplit.Schedule(tile_nnz=T, tile_mode=1)
    plan = rowsplit.RowSplitPlan(torch.from_numpy(ip), n, torch.float32, sch, 0)
    R = plan.tile_start.numpy()
    E = plan.tile_entry.numpy()
    assert R[0] == 0 and R[-1] == n and E[0] == 0 and E[-1] == nnz
    assert np.all(np.diff(R) >= 0) and np.all(np.diff(E) >= 0)
    assert np.all(np.diff(E) <= T) and np.all(np.diff(R) <= rowsplit.RowSplitPlan.ENTRY_ROW_CAP)
    assert set((np.arange(1, -(-nnz // T)) * T).tolist()) <= set(E.tolist())
    for t in range(plan.n_tiles):
        rows = np.arange(R[t], R[t + 1])
        if rows.size == 0:
            continue
        assert np.all(ip[rows] >= E[t]) and np.all(ip[rows] <= E[t + 1])
        nonempty = lens[rows] > 0
        assert np.all(ip[rows][nonempty] < E[t + 1])
        assert np.all(ip[rows[:-1] + 1] <= E[t + 1])
    assert plan.n_long == 0


def test_entry_tile_empty_matrix():
    ip = torch.zeros(7, dtype=torch.int64)
    plan = rowsplit.RowSplitPlan(ip, 6, torch.float64, rowsplit.Schedule(tile_nnz=2048, tile_mode=1), 0)
    assert plan.n_tiles == 1 and plan.tile_start.tolist() == [0, 6]
    assert plan.tile_entry.tolist() == [0, 0]


def test_sliced_layout_structure():
    """The slice-major copy: compact rows ordered by (slice, row), every entry
    in its slice, the compact-row sum over rows equal to A x (fp64, SciPy),
    combine bounds selecting (slice, 256-row block) ranges, and the XCD group
    ranges covering every tile once."""
    import scipy.sparse as sp
    from flagsparse_public_amd import slicing
    ip, cols, vals = synthetic.powerlaw_csr(20000, 11, alpha=1.6, max_len=30000)
    n = ip.size - 1
    s = slicing.SlicedCsr(torch.from_numpy(vals), torch.from_numpy(cols), torch.from_numpy(ip),
                          n, n, 3000, rowsplit.Schedule(), 0)
    assert s.n_slices == 8 and s.width == 2500
    x = np.random.default_rng(0).uniform(-1, 1, n)
    cp = s.cptr.numpy().astype(np.int64)
    cr = s.compact_row.numpy()
    sc = s.cols.numpy().astype(np.int64)
    sv = s.vals.numpy().astype(np.float64)
    z = np.add.reduceat(sv * x[sc], cp[:-1]) if s.n_compact else np.zeros(0)
    y = np.zeros(n)
    np.add.at(y, cr, z)
    ref = sp.csr_matrix((vals.astype(np.float64), cols, ip), shape=(n, n)) @ x
    assert np.abs(y - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    sl = sc[cp[:-1]] // s.width
    key = sl * n + cr
    assert np.all(np.diff(key) > 0)
    assert np.all(np.repeat(sl, np.diff(cp)) == sc // s.width)
    b = s.bounds.numpy().reshape(s.n_slices, -1)
    for k in range(s.n_slices):
        for bb in range(b.shape[1] - 1):
            seg = np.arange(b[k, bb], b[k, bb + 1])
            assert seg.size <= slicing.COMBINE_ROWS
            assert np.all(sl[seg] == k) and np.all(cr[seg] // slicing.COMBINE_ROWS == bb)
    gt = s.group_tiles.tolist()
    assert gt[0] == 0 and gt[-1] == s.plan.n_tiles and gt == sorted(gt)
    assert s.group_blocks == max(np.diff(gt))
    # row masks: bit set exactly for the (slice, row) pairs with a compact
    # row, and the popcount rule finds each compact row's index
    nb = b.shape[1] - 1
    words = s.row_mask.numpy().view(np.uint32).reshape(s.n_slices, nb, 8)
    bits = ((words[..., None] >> np.arange(32, dtype=np.uint32)) & 1).reshape(s.n_slices, nb * 256)
    present = np.zeros((s.n_slices, nb * 256), np.uint8)
    present[sl, cr] = 1
    assert np.array_equal(bits, present)
    for i in np.random.default_rng(1).choice(s.n_compact, 2000, replace=False):
        k, r = sl[i], cr[i]
        bb = r // slicing.COMBINE_ROWS
        rank = int(bits[k, bb * 256:r].sum())
        assert b[k, bb] + rank == i


def test_synthetic_generators_deterministic_and_sorted():
    a = synthetic.poisson_csr(2000, 12.0, 9)
    b = synthetic.poisson_csr(2000, 12.0, 9)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    ip, cols, vals = a
    lens = np.diff(ip)
    rows = np.repeat(np.arange(lens.size), lens)
    key = rows * 2000 + cols.astype(np.int64)
    assert np.all(np.diff(key) > 0)  # sorted, no duplicates within a row
    assert lens.min() >= 1 and np.all(np.abs(vals) <= 1)
    ip, cols, _ = synthetic.banded_random_csr(5000, 3)
    lens = np.diff(ip)
    assert lens.max() == 15 and lens.min() == 12
    rows = np.repeat(np.arange(5000), lens)
    band = np.abs(cols.astype(np.int64) - rows) <= 3
    # band entries per row equal the clipped band width; the rest are off-band
    assert np.array_equal(np.bincount(rows, weights=band, minlength=5000).astype(int),
                          lens - 8)
    ip, cols, _ = synthetic.block_diag_csr(20, 0)
    assert np.array_equal(np.diff(ip), [8] * 16 + [4] * 4)
    ip, cols, _ = synthetic.banded_csr(10, 0, 2)
    assert np.array_equal(np.diff(ip), [3, 4, 5, 5, 5, 5, 5, 5, 4, 3])


def test_powerlaw_lengths_distribution():
    rng = np.random.Generator(np.random.PCG64(2))
    l = synthetic.powerlaw_lengths(rng, 200000, 2.0, 1, 100000)
    # continuous density ~ l^-2 on [1, 1e5], ceiling: mean = ln(1e5)/(1-1e-5) + ~0.577
    assert abs(l.mean() - (np.log(1e5) / (1 - 1e-5) + 0.577)) < 0.6
    assert l.min() >= 1 and l.max() <= 100000
    # P(len > 128) = (1/128 - 1/1e5) / (1 - 1/1e5) for the ceiling of l > 128
    assert abs((l > 128).mean() - (1 / 128 - 1e-5)) < 1e-3


def test_batched_cuts_and_items():
    from flagsparse_public_amd.batched import BatchedCsr, _cuts
    lens = torch.tensor([3, 0, 5000, 2, 2, 9000, 1, 0, 4])
    ip = torch.zeros(lens.numel() + 1, dtype=torch.int64)
    ip[1:] = torch.cumsum(lens, 0)
    cuts = _cuts(ip, 1.0, 100, 3, long_row=4096)
    c = cuts.tolist()
    assert c[0] == 0 and c[-1] == lens.numel() and c == sorted(set(c))
    # both long rows (2 and 5) are isolated into single-row items
    for r in (2, 5):
        assert r in c and r + 1 in c
    # no item spans more than 3 rows
    assert max(b - a for a, b in zip(c[:-1], c[1:])) <= 3
    items = BatchedCsr._items(7, 1, cuts)
    assert items.shape[1] == 3 and (items[:, 0] == (7 | (1 << 32))).all()
    assert items[:, 1].tolist() == c[:-1] and items[:, 2].tolist() == c[1:]


def test_batch_pack_validates_on_host():
    import ctypes
    from flagsparse_public_amd import _lib
    from flagsparse_public_amd.batched import FsBatchEntry, _lib_batch
    lib = _lib_batch()
    assert lib.fs_batch_matrix_bytes() >= 8 * 11
    ents = (FsBatchEntry * 2)()
    for e in ents:
        e.A = _lib.FsCsr(value_type=_lib.FS_F32, index_type=_lib.FS_I32, ptr_type=_lib.FS_I32,
                         reserved=0, n_rows=4, n_cols=4, nnz=4, row_ptr=64, col_idx=128, values=256)
        e.n_dense, e.ldb, e.ldc, e.group_lanes = 8, 8, 8, 4
    buf = (ctypes.c_uint8 * (lib.fs_batch_matrix_bytes() * 2))()
    assert lib.fs_batch_pack(ents, 2, ctypes.addressof(buf)) == _lib.FS_OK
    ents[1].group_lanes = 3
    assert lib.fs_batch_pack(ents, 2, ctypes.addressof(buf)) == _lib.FS_ERR_VALUE
    ents[1].group_lanes = 4
    ents[1].A.value_type = _lib.FS_F64
    assert lib.fs_batch_pack(ents, 2, ctypes.addressof(buf)) == _lib.FS_ERR_TYPE


def test_batched_cuts_small_matrices():
    from flagsparse_public_amd.batched import _cuts
    ip = torch.tensor([0, 2, 3, 3])
    assert _cuts(ip, 12.0, 96 * 1024, 2048, 4096).tolist() == [0, 3]
    assert _cuts(torch.tensor([0]), 12.0, 100, 4).tolist() == [0]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_spmm_candidate_buckets_match_length_ranges(dtype):
    """The diagnose candidate buckets (spmm_csr.py:1835-1873): every row with
    at most 2048 entries lands in the one spec range holding its length (the
    first range also takes empty rows), rows over 2048 go to "split", labels
    and kinds follow the dtype's table."""
    from types import SimpleNamespace
    from flagsparse_public_amd.sparse_operations import spmm_csr as m
    rng = np.random.default_rng(4)
    lens = np.concatenate([rng.integers(0, 2200, 3000), [0, 16, 17, 32, 33, 2048, 2049, 5000]])
    prepared = SimpleNamespace(row_lengths=torch.from_numpy(lens.astype(np.int64)),
                               data=torch.empty(0, dtype=dtype))
    buckets = m._build_spmm_opt_candidate_buckets(prepared)
    specs = m._select_spmm_opt_candidate_specs(dtype)
    seen = np.zeros(lens.size, np.int64)
    for b in buckets:
        rows = b["rows"].numpy()
        assert b["rows"].dtype == torch.int32
        seen[rows] += 1
        if b["label"] == "split":
            assert b["kind"] == "split" and np.all(lens[rows] > 2048)
            assert set(rows) == set(np.nonzero(lens > 2048)[0])
            continue
        spec = next(sp for sp in specs if sp["label"] == b["label"])
        lo, hi = spec["min_row_nnz"], spec["max_row_nnz"]
        want = np.nonzero((lens <= hi) & ((lens >= lo) if lo > 0 else True))[0]
        assert np.array_equal(np.sort(rows), want)
        assert b["kind"] == spec["kind"] and b["block_nnz"] == spec["block_nnz"]
        assert b.get("segments") == spec.get("segments")
    assert np.all(seen == 1)
    kinds = {sp["label"]: sp["kind"] for sp in specs}
    assert kinds["short_16"] == ("segmented" if dtype == torch.float64 else "micro")


def test_supported_spmv_schedule():
    """Library-chosen entry tiles of 1024 / 4096 run as 2048 where only 2048
    is built (int64 columns, values other than fp32 / fp64); x-window
    schedules keep their tiles for fp32 / fp64 with any column type."""
    import torch
    from flagsparse_public_amd.rowsplit import Schedule, supported_spmv_schedule
    t4 = Schedule(tile_nnz=4096, tile_mode=1)
    xw = Schedule(tile_nnz=4096, tile_mode=1, x_window=-1)
    rows = Schedule(group_lanes=16)
    assert supported_spmv_schedule(t4, torch.float32, torch.int32) == t4
    assert supported_spmv_schedule(t4, torch.float64, torch.int64).tile_nnz == 2048
    assert supported_spmv_schedule(t4, torch.float16, torch.int32).tile_nnz == 2048
    assert supported_spmv_schedule(xw, torch.float32, torch.int64) == xw
    assert supported_spmv_schedule(xw, torch.float32, torch.int64, 2**31 - 1) == xw
    # beyond the x-window layout's reach the launch runs plain entry tiles
    assert supported_spmv_schedule(xw, torch.float32, torch.int64, 2**31).tile_nnz == 2048
    assert supported_spmv_schedule(xw, torch.float32, torch.int32, 2**31) == xw
    assert supported_spmv_schedule(xw, torch.complex64, torch.int32).tile_nnz == 2048
    assert supported_spmv_schedule(rows, torch.float16, torch.int64) == rows
    sliced = Schedule(tile_nnz=1024, tile_mode=1, x_slice_cols=1 << 18)
    out = supported_spmv_schedule(sliced, torch.float32, torch.int64)
    assert out.tile_nnz == 2048 and out.x_slice_cols == 1 << 18


def test_stable_order_narrow_keys():
    """Stable orders of keys narrowed to 16 / 32 bits equal the int64 order,
    equal keys in input order (every bound class, duplicates included)."""
    from flagsparse_public_amd.sortkey import narrow_key, stable_order
    g = torch.Generator().manual_seed(9)
    for bound in (7, 2**15, 2**15 + 1, 2**31, 2**31 + 5):
        key = torch.randint(0, min(bound, 2**40), (20000,), generator=g, dtype=torch.int64)
        key[::7] = bound - 1
        want = torch.argsort(key, stable=True)
        assert torch.equal(stable_order(key, bound), want), bound
        nk = narrow_key(key, bound)
        assert nk.dtype == (torch.int16 if bound <= 2**15 else
                            torch.int32 if bound <= 2**31 else torch.int64)
