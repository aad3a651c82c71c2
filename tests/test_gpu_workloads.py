"""Full-size GPU checks of the batched and learned-selection workloads.

* The uneven batch at its real size: 512 square matrices (n log-uniform in
  [1e3, 2e5], Poisson row-length mean U[4, 64]), each with an SpMV and an
  8-column SpMM, run as one persistent launch replayed from a HIP graph.
  Every replay equals the eager launch bit for bit; a spread of matrices
  equals the single-matrix launches bit for bit and matches the fp64 oracle.
* Learned selection on held-out matrices (seeds 7000.. and the large
  supplement 9500..): ``prepare_spmv_csr`` / ``prepare_spmm_csr_opt`` run
  the whole selection on the device; the features equal the ones recorded
  in the committed timing database, the choice equals the oracle's MLP
  forward on them, and the chosen kernel matches the oracle on a sample of
  rows (every long row included).
"""

import json
import os

import numpy as np
import pytest
import torch

from tests._util import host64, row_rel_err

pytestmark = pytest.mark.gpu

from flagsparse_public_amd import (_lib, flagsparse_spmm_csr_opt, flagsparse_spmv_csr,  # noqa: E402
                                   prepare_spmm_csr_opt, prepare_spmv_csr, synthetic)
from flagsparse_public_amd.autotune import TimingDB, default_db_dir  # noqa: E402
from flagsparse_public_amd.batched import BatchedCsr  # noqa: E402
from flagsparse_public_amd.batched_workload import uneven_batch  # noqa: E402
from flagsparse_public_amd.rowsplit import Schedule  # noqa: E402

DEV = torch.device("cuda", 0)


def _host_csr(m):
    return (m["indptr"].cpu().numpy().astype(np.int64), m["indices"].cpu().numpy(),
            m["data"].cpu().numpy())


def test_uneven_batch_512_graph_replay(orc):
    mats, meta = uneven_batch(512, 4, DEV)
    print(f"[c4] {meta['count']} matrices, {meta['rows']} rows, {meta['nnz']} entries",
          flush=True)
    assert meta["count"] == 512 and meta["nnz"] > 500_000_000
    bat = BatchedCsr(mats)
    bat.run()
    torch.cuda.synchronize()
    ys = [y.clone() for y in bat.y]
    Cs = [C.clone() for C in bat.C]
    g = bat.capture()
    for rep in range(3):
        for y in bat.y:
            y.fill_(float("nan"))  # eager work between replays
        for C in bat.C:
            C.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert all(torch.equal(a, b) for a, b in zip(bat.y, ys)), rep
        assert all(torch.equal(a, b) for a, b in zip(bat.C, Cs)), rep
    assert int(bat.queue.sum().item()) == 0
    spread = sorted(set([0, 1, 511] + list(range(37, 512, 73))
                        + [int(np.argmax([m["shape"][0] for m in mats]))]))
    for i in spread:
        m = mats[i]
        ip, cols, vals = _host_csr(m)
        x = m["x"].cpu().numpy()
        y64 = orc.spmv(ip, cols, vals, x)
        cond = orc.spmv_abs(ip, cols, vals, x)
        assert row_rel_err(host64(bat.y[i]), y64, cond).max() <= 1e-5, i
        B = m["B"].cpu().numpy()
        C64 = orc.spmm(ip, cols, vals, B)
        S = orc.spmm_abs(ip, cols, vals, B)
        assert (np.abs(host64(bat.C[i]) - C64) / np.where(S > 0, S, 1)).max() <= 1e-5, i
    # the same rows as the single-matrix launches with the same lane groups
    for i in spread[:4]:
        m = mats[i]
        lanes = int(bat.group_lanes[i])
        pv = prepare_spmv_csr(m["data"], m["indices"], m["indptr"], m["shape"],
                              schedule=Schedule(group_lanes=lanes, short_max=1 << 40,
                                                chunk_nnz=512))
        assert torch.equal(bat.y[i], flagsparse_spmv_csr(x=m["x"], prepared=pv, use_opt=True)), i
        pm = prepare_spmm_csr_opt(m["data"], m["indices"], m["indptr"], m["shape"],
                                  schedule=Schedule(short_max=1 << 40, chunk_nnz=256))
        assert torch.equal(bat.C[i], flagsparse_spmm_csr_opt(B=m["B"], prepared=pm)), i


# --------------------------------------------------------------- selection

DB_DIR = default_db_dir()
HELD_OUT = [7000, 7001, 7002, 7003, 7004, 7006, 7008, 7012, 7013, 7016, 9500, 9504]


def _load(op, db_name=None, model_name=None):
    db_path = os.path.join(DB_DIR, db_name or f"{op}_test_db.npz")
    model_path = os.path.join(DB_DIR, model_name or f"{op}_f32_mlp.json")
    if not (os.path.exists(db_path) and os.path.exists(model_path)):
        pytest.skip(f"no committed selector for {op}")
    db = TimingDB.load(db_path)
    with open(model_path) as fh:
        model = json.load(fh)
    return db, model


def _oracle_scores(orc, model, feats):
    mu = np.asarray(model["mu"], np.float32)
    sd = np.asarray(model["sd"], np.float32)
    Z = ((np.asarray(feats, np.float32) - mu) / sd).astype(np.float64)[None, :]
    p = np.asarray(model["params"], np.float64)
    i, h, o = model["in_dim"], model["hidden"], model["out_dim"]
    W, k = [], 0
    for s in [(i, h), (h,), (h, h), (h,), (h, o), (o,)]:
        n = int(np.prod(s))
        W.append(p[k:k + n].reshape(s))
        k += n
    return orc.mlp_forward(Z, *W)[2][0]


def _row_sample(indptr, rng, k=3000):
    lens = np.diff(indptr)
    rows = rng.choice(lens.size, min(k, lens.size), replace=False)
    longest = np.argsort(lens)[-64:]
    return np.unique(np.concatenate([rows, longest, [0, lens.size - 1]]))


def _sub_csr(ip, cols, vals, rows):
    lens = ip[rows + 1] - ip[rows]
    sub_ip = np.zeros(rows.size + 1, np.int64)
    np.cumsum(lens, out=sub_ip[1:])
    idx = np.concatenate([np.arange(ip[r], ip[r + 1]) for r in rows]) if sub_ip[-1] else \
        np.zeros(0, np.int64)
    return sub_ip, cols[idx], vals[idx], idx


def _check_selection(orc, op, seed, p, db, model):
    ids = list(db.ids)
    if seed not in ids:
        pytest.skip(f"seed {seed} not in the committed {op} database")
    row = ids.index(seed)
    how = p.schedule_record["chosen_by"]
    print(f"[c5] {op} seed {seed}: {how}", flush=True)
    stats = _lib.analyze_device(p.csr_struct(), DEV)
    feats = _lib.features(stats, p.n_rows, p.n_cols, p.data.numel(), DEV).double().cpu().numpy()
    rec = db.features[row]
    assert np.allclose(feats, rec, rtol=1e-6, atol=1e-6), (feats, rec)
    if not how.startswith("learned:"):
        # only matrices beyond the trained sizes fall back to the fixed rule
        assert how.startswith("default (outside")
        return None
    idx = int(how.split(":")[1])
    scores = _oracle_scores(orc, model, rec)
    best = int(np.argmin(scores))
    assert idx == best or abs(scores[idx] - scores[best]) <= 1e-4, (idx, best, scores)
    t = db.times[row]
    print(f"[c5] {op} seed {seed}: chosen {idx}, regret {t[idx] / t.min():.3f}", flush=True)
    return idx


@pytest.mark.parametrize("seed", HELD_OUT)
def test_learned_spmv_on_held_out(orc, seed):
    db, model = _load("spmv")
    family, n = synthetic.selector_matrix_spec(seed)
    indptr, cols, vals, _ = synthetic.family_csr_torch(family, n, seed, DEV)
    p = prepare_spmv_csr(vals, cols, indptr, (n, n))
    _check_selection(orc, "spmv", seed, p, db, model)
    g = torch.Generator(device=DEV)
    g.manual_seed(seed + 1)
    x = torch.rand(n, generator=g, device=DEV) * 2 - 1
    y = flagsparse_spmv_csr(x=x, prepared=p, use_opt=True)
    ip, ch, vh = indptr.cpu().numpy(), cols.cpu().numpy(), vals.cpu().numpy()
    rows = _row_sample(ip, np.random.default_rng(seed))
    sip, sc, sv, _ = _sub_csr(ip, ch, vh, rows)
    xh = x.cpu().numpy()
    y64 = orc.spmv(sip, sc, sv, xh)
    cond = orc.spmv_abs(sip, sc, sv, xh)
    assert row_rel_err(host64(y)[rows], y64, cond).max() <= 1e-5


@pytest.mark.parametrize("seed", [7000, 7002, 7013, 9500])
def test_learned_spmv_fp64_on_held_out(orc, seed):
    """fp64 values: the model trained on fp64 timings (the fp32-trained one
    where none is packaged) chooses from the structure features recorded in
    the fp64 held-out database, and the chosen kernel meets the fp64 bar on a
    sample of rows."""
    name = "spmv_f64_mlp.json" if os.path.exists(os.path.join(DB_DIR, "spmv_f64_mlp.json")) \
        else "spmv_f32_mlp.json"
    db, model = _load("spmv", "spmv_f64_test_db.npz", name)
    family, n = synthetic.selector_matrix_spec(seed)
    indptr, cols, vals, _ = synthetic.family_csr_torch(family, n, seed, DEV)
    vals = vals.double()
    p = prepare_spmv_csr(vals, cols, indptr, (n, n))
    assert p.schedule_record["chosen_by"].startswith(("learned:", "default (outside"))
    _check_selection(orc, "spmv", seed, p, db, model)
    g = torch.Generator(device=DEV)
    g.manual_seed(seed + 1)
    x = (torch.rand(n, generator=g, device=DEV) * 2 - 1).double()
    y = flagsparse_spmv_csr(x=x, prepared=p, use_opt=True)
    ip, ch, vh = indptr.cpu().numpy(), cols.cpu().numpy(), vals.cpu().numpy()
    rows = _row_sample(ip, np.random.default_rng(seed))
    sip, sc, sv, _ = _sub_csr(ip, ch, vh, rows)
    xh = x.cpu().numpy()
    y64 = orc.spmv(sip, sc, sv, xh)
    cond = orc.spmv_abs(sip, sc, sv, xh)
    assert row_rel_err(host64(y)[rows], y64, cond).max() <= 1e-12


@pytest.mark.parametrize("seed", HELD_OUT[:8] + HELD_OUT[10:11])
def test_learned_spmm64_on_held_out(orc, seed):
    db, model = _load("spmm64")
    family, n = synthetic.selector_matrix_spec(seed)
    indptr, cols, vals, _ = synthetic.family_csr_torch(family, n, seed, DEV)
    p = prepare_spmm_csr_opt(vals, cols, indptr, (n, n))
    _check_selection(orc, "spmm64", seed, p, db, model)
    g = torch.Generator(device=DEV)
    g.manual_seed(seed + 2)
    B = torch.rand((n, 64), generator=g, device=DEV) * 2 - 1
    C = flagsparse_spmm_csr_opt(B=B, prepared=p)
    ip, ch, vh = indptr.cpu().numpy(), cols.cpu().numpy(), vals.cpu().numpy()
    rows = _row_sample(ip, np.random.default_rng(seed), k=1500)
    sip, sc, sv, _ = _sub_csr(ip, ch, vh, rows)
    # the B rows the sampled entries gather, entry by entry (same sum order)
    Bg = B[torch.from_numpy(sc.astype(np.int64)).to(DEV)].double().cpu().numpy()
    pos = np.arange(sc.size)
    C64 = orc.spmm(sip, pos, sv, Bg)
    S = orc.spmm_abs(sip, pos, sv, Bg)
    got = host64(C[torch.from_numpy(rows).to(DEV)])
    assert (np.abs(got - C64) / np.where(S > 0, S, 1)).max() <= 1e-5
