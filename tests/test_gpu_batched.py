"""GPU tests of the batched persistent path (MI355X).

* Every batched row equals the single-matrix launch with the same lane-group
  width bit for bit (SpMV and SpMM), on uneven matrices with empty rows.
* A sample of matrices matches the fp64 oracle (1e-5 * sum|a x| per row).
* Graph replays: the captured launch, replayed several times with eager work
  and synchronisations between replays, reproduces the eager result each time.
* A single very long row goes through the block-wide reduction.
"""

import numpy as np
import pytest
import torch

from tests._util import host64, row_rel_err

pytestmark = pytest.mark.gpu

from flagsparse_public_amd import flagsparse_spmm_csr_opt, flagsparse_spmv_csr  # noqa: E402
from flagsparse_public_amd import prepare_spmm_csr_opt, prepare_spmv_csr, synthetic  # noqa: E402
from flagsparse_public_amd.batched import BatchedCsr  # noqa: E402
from flagsparse_public_amd.rowsplit import Schedule  # noqa: E402

DEV = torch.device("cuda", 0)


def _batch_mats(count, seed, n_dense=8, max_n=20000):
    rng = np.random.Generator(np.random.PCG64(seed))
    mats = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(50), np.log(max_n))))
        mean = float(rng.uniform(0.5, 40))
        lens = np.minimum(rng.poisson(mean, n), n)
        ip = np.zeros(n + 1, np.int64)
        np.cumsum(lens, out=ip[1:])
        cols = rng.integers(0, n, ip[-1]).astype(np.int32)
        vals = rng.uniform(-1, 1, ip[-1]).astype(np.float32)
        x = rng.uniform(-1, 1, n).astype(np.float32)
        B = rng.uniform(-1, 1, (n, n_dense)).astype(np.float32)
        mats.append(dict(ip=ip, cols=cols, vals=vals, x=x, B=B, n=n,
                         group_lanes=int(2 ** rng.integers(0, 7))))
    return mats


def _to_dev(mats):
    out = []
    for m in mats:
        out.append(dict(data=torch.from_numpy(m["vals"]).to(DEV),
                        indices=torch.from_numpy(m["cols"]).to(DEV),
                        indptr=torch.from_numpy(m["ip"]).to(DEV), shape=(m["n"], m["n"]),
                        x=torch.from_numpy(m["x"]).to(DEV), B=torch.from_numpy(m["B"]).to(DEV),
                        group_lanes=m["group_lanes"]))
    return out


def test_batched_matches_single_launches_bitwise():
    mats = _batch_mats(40, 1)
    dm = _to_dev(mats)
    bat = BatchedCsr(dm)
    bat.run()
    torch.cuda.synchronize()
    for i, m in enumerate(dm):
        sch = Schedule(group_lanes=m["group_lanes"], short_max=1 << 40, chunk_nnz=512)
        pv = prepare_spmv_csr(m["data"], m["indices"], m["indptr"], m["shape"], schedule=sch)
        y = flagsparse_spmv_csr(x=m["x"], prepared=pv, use_opt=True)
        assert torch.equal(bat.y[i], y), i
        pm = prepare_spmm_csr_opt(m["data"], m["indices"], m["indptr"], m["shape"],
                                  schedule=Schedule(short_max=1 << 40, chunk_nnz=256))
        C = flagsparse_spmm_csr_opt(B=m["B"], prepared=pm)
        assert torch.equal(bat.C[i], C), i


def test_batched_against_oracle(orc):
    mats = _batch_mats(12, 2, n_dense=8, max_n=5000)
    bat = BatchedCsr(_to_dev(mats))
    bat.run()
    torch.cuda.synchronize()
    for i, m in enumerate(mats):
        y64 = orc.spmv(m["ip"], m["cols"], m["vals"], m["x"])
        cond = orc.spmv_abs(m["ip"], m["cols"], m["vals"], m["x"])
        assert row_rel_err(host64(bat.y[i]), y64, cond).max() <= 1e-5
        C64 = orc.spmm(m["ip"], m["cols"], m["vals"], m["B"])
        S = orc.spmm_abs(m["ip"], m["cols"], m["vals"], m["B"])
        assert (np.abs(host64(bat.C[i]) - C64) / np.where(S > 0, S, 1)).max() <= 1e-5


def test_batched_graph_replays_match_eager():
    mats = _batch_mats(30, 3)
    bat = BatchedCsr(_to_dev(mats))
    bat.run()
    torch.cuda.synchronize()
    ys = [y.clone() for y in bat.y]
    Cs = [C.clone() for C in bat.C]
    g = bat.capture()
    for rep in range(4):
        for y in bat.y:
            y.fill_(float("nan"))  # eager work between replays
        for C in bat.C:
            C.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(bat.y, ys):
            assert torch.equal(a, b), rep
        for a, b in zip(bat.C, Cs):
            assert torch.equal(a, b), rep
    assert int(bat.queue.sum().item()) == 0  # self-reset after every launch


def test_batched_long_row_block_reduction(orc):
    rng = np.random.default_rng(4)
    n, m_cols = 6, 200000
    lens = np.array([150000, 3, 0, 9000, 1, 5000])
    ip = np.zeros(n + 1, np.int64)
    np.cumsum(lens, out=ip[1:])
    cols = np.concatenate([np.sort(rng.choice(m_cols, l, replace=False)) for l in lens]).astype(np.int32)
    vals = rng.uniform(-1, 1, ip[-1]).astype(np.float32)
    x = rng.uniform(-1, 1, m_cols).astype(np.float32)
    d = dict(data=torch.from_numpy(vals).to(DEV), indices=torch.from_numpy(cols).to(DEV),
             indptr=torch.from_numpy(ip).to(DEV), shape=(n, m_cols), x=torch.from_numpy(x).to(DEV))
    bat = BatchedCsr([d], long_row=4096)
    bat.run()
    y64 = orc.spmv(ip, cols, vals, x)
    cond = orc.spmv_abs(ip, cols, vals, x)
    assert row_rel_err(host64(bat.y[0]), y64, cond).max() <= 1e-6


def test_batched_uneven_workload_sample(orc):
    """A slice of the 512-matrix uneven workload (n log-uniform in
    [1e3, 2e5], Poisson mean U[4, 64], SpMV + 8-column SpMM)."""
    from flagsparse_public_amd.batched_workload import uneven_batch
    mats, meta = uneven_batch(count=24, seed=4, device=DEV)
    bat = BatchedCsr(mats)
    bat.run()
    torch.cuda.synchronize()
    for i in (0, 7, 23):
        m = mats[i]
        ip = m["indptr"].cpu().numpy()
        cols = m["indices"].cpu().numpy()
        vals = m["data"].cpu().numpy()
        x = m["x"].cpu().numpy()
        y64 = orc.spmv(ip, cols, vals, x)
        cond = orc.spmv_abs(ip, cols, vals, x)
        assert row_rel_err(host64(bat.y[i]), y64, cond).max() <= 1e-5
