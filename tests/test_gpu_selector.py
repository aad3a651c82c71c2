"""GPU tests of the learned selector's HIP kernels (MI355X).

* fs_csr_features equals the oracle's 16 statistics (fp32 rounding of fp64).
* fs_mlp_forward / fs_mlp_backward equal the fp64 oracle within fp32
  rounding (sums are carried in fp64 on the device), and the device gradient
  is deterministic.
* Adam matches the oracle's update; training lowers the loss.
* fs_mlp_select picks the oracle forward's argmin.
"""

import numpy as np
import pytest
import torch

from tests._util import to_dev

pytestmark = pytest.mark.gpu

from flagsparse_public_amd import _lib, synthetic  # noqa: E402
from flagsparse_public_amd.selector import Mlp, targets_from_times  # noqa: E402
from flagsparse_public_amd.sparse_operations.spmv_csr import prepare_spmv_csr  # noqa: E402

DEV = torch.device("cuda", 0)


@pytest.mark.parametrize("case", ["poisson", "powerlaw", "empty"])
def test_features_match_oracle(orc, case):
    if case == "poisson":
        ip, cols, vals = synthetic.poisson_csr(5000, 9.0, 1, lo=0)
    elif case == "powerlaw":
        ip, cols, vals = synthetic.powerlaw_csr(8000, 2, alpha=1.7, max_len=6000)
    else:
        ip, cols, vals = np.zeros(4, np.int64), np.zeros(0, np.int32), np.zeros(0, np.float32)
    n = ip.size - 1
    m = max(n, 1)
    p = prepare_spmv_csr(to_dev(vals), to_dev(cols), to_dev(ip), (n, m))
    stats = _lib.analyze_device(p.csr_struct(), DEV)
    f = _lib.features(stats, n, m, vals.size, DEV).cpu().numpy()
    want = orc.row_features(ip, m).astype(np.float32)
    assert np.allclose(f, want, rtol=1e-6, atol=1e-7)


def _rand_problem(seed, batch=37, out_dim=11):
    rng = np.random.default_rng(seed)
    mlp = Mlp(out_dim, DEV, seed=seed)
    X = rng.normal(size=(batch, 16)).astype(np.float32)
    T = rng.normal(size=(batch, out_dim)).astype(np.float32)
    M = (rng.random((batch, out_dim)) < 0.85).astype(np.float32)
    return mlp, X, T, M


def test_forward_backward_match_oracle(orc):
    mlp, X, T, M = _rand_problem(3)
    P = mlp.unpack()
    H1, H2, Y = mlp.forward(to_dev(X))
    oH1, oH2, oY = orc.mlp_forward(X.astype(np.float64), *P)
    assert np.allclose(Y.cpu().numpy(), oY, rtol=1e-5, atol=1e-6)
    assert np.allclose(H1.cpu().numpy(), oH1, rtol=1e-5, atol=1e-6)
    loss, grad = mlp.loss_and_grad(to_dev(X), to_dev(T), to_dev(M))
    oloss, og = orc.mlp_backward(X.astype(np.float64), P[2], P[4], oH1, oH2, oY,
                                 T.astype(np.float64), M.astype(np.float64))
    g = grad.double().cpu().numpy()
    og_flat = np.concatenate([a.ravel() for a in og])
    assert float(loss.item()) == pytest.approx(oloss, rel=1e-5)
    scale = np.abs(og_flat).max()
    assert np.max(np.abs(g - og_flat)) <= 1e-5 * scale
    _, grad2 = mlp.loss_and_grad(to_dev(X), to_dev(T), to_dev(M))
    assert torch.equal(grad, grad2)


def test_adam_matches_oracle(orc):
    mlp, X, T, M = _rand_problem(5)
    p0 = mlp.params.double().cpu().numpy()
    _, grad = mlp.loss_and_grad(to_dev(X), to_dev(T), to_dev(M))
    g = grad.double().cpu().numpy()
    mlp.adam(grad, lr=1e-2)
    p = p0.copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    orc.adam_step(p, g, m, v, 1e-2, 0.9, 0.999, 1e-8, 1)
    assert np.allclose(mlp.params.double().cpu().numpy(), p, rtol=1e-6, atol=1e-7)


def test_training_reduces_loss():
    rng = np.random.default_rng(8)
    feats = rng.normal(size=(300, 16))
    # times with a learnable structure: variant k is fast where feature k%16 is large
    base = np.exp(rng.normal(size=(300, 6)) * 0.05)
    for k in range(6):
        base[:, k] *= np.exp(-0.5 * np.tanh(feats[:, k]))
    T, M = targets_from_times(base)
    mlp = Mlp(6, DEV, seed=1)
    Z = torch.from_numpy(((feats - feats.mean(0)) / feats.std(0)).astype(np.float32)).to(DEV)
    losses = mlp.fit(Z, torch.from_numpy(T).to(DEV), torch.from_numpy(M).to(DEV), epochs=600,
                     lr=3e-3, log_every=100)
    assert losses[-1][1] < 0.25 * losses[0][1]


def test_select_kernel_is_oracle_argmin(orc):
    mlp, X, _, _ = _rand_problem(9, batch=20, out_dim=16)
    mu = np.zeros(16, np.float32)
    sd = np.ones(16, np.float32)
    P = mlp.unpack()
    _, _, oY = orc.mlp_forward(X.astype(np.float64), *P)
    for i in range(X.shape[0]):
        choice = torch.empty(1, dtype=torch.int32, device=DEV)
        scores = torch.empty(16, dtype=torch.float32, device=DEV)
        _lib.mlp_select(mlp.net, to_dev(X[i]), to_dev(mu), to_dev(sd), choice, scores, DEV)
        assert np.allclose(scores.cpu().numpy(), oY[i], rtol=1e-5, atol=1e-6)
        top2 = np.sort(oY[i])[:2]
        if top2[1] - top2[0] > 1e-4:
            assert int(choice.item()) == int(np.argmin(oY[i]))
