"""Learned schedule selection: a 16 -> 64 -> 64 -> K perceptron, trained and
run on the GPU by the library's own HIP kernels (csrc/selector.hip).

Inputs are the 16 row-length statistics of ``fs_csr_features``, standardised
with the training set's mean / std.  Outputs are predicted log2 of each
variant's time over the fastest variant's time (trained with masked MSE);
the selector picks the smallest output.  At prepare time the whole chain --
analysis pass, features, standardisation, forward, argmin -- runs on the
device, and the host reads back one small buffer.

Models are JSON files (weights, normalisation, variant list version, device
name) under ``selector_models/``.
"""

from __future__ import annotations

import json
import math
import os
import warnings

import numpy as np
import torch

from . import _lib
from .autotune import SPMM_VARIANTS, SPMV_VARIANTS, VARIANTS_VERSION, default_db_dir

IN_DIM = 16
HIDDEN = 64


class Mlp:
    """Device-resident perceptron parameters with Adam state."""

    def __init__(self, out_dim: int, device, hidden: int = HIDDEN, in_dim: int = IN_DIM, seed: int = 0,
                 params: np.ndarray | None = None):
        self.in_dim, self.hidden, self.out_dim = in_dim, hidden, out_dim
        self.device = torch.device(device)
        n = _lib.mlp_param_count(in_dim, hidden, out_dim)
        if params is None:
            params = self.init_params(in_dim, hidden, out_dim, seed)
        params = np.asarray(params, np.float32)
        if params.size != n:
            raise ValueError(f"expected {n} parameters, got {params.size}")
        self.params = torch.from_numpy(params.copy()).to(self.device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        self.net = _lib.mlp_struct(self.params, in_dim, hidden, out_dim)

    @staticmethod
    def init_params(in_dim, hidden, out_dim, seed):
        """He-uniform weights, zero biases, in the packed fs_mlp layout."""
        rng = np.random.Generator(np.random.PCG64(seed))

        def w(fan_in, fan_out):
            lim = math.sqrt(6.0 / fan_in)
            return rng.uniform(-lim, lim, (fan_in, fan_out))

        parts = [w(in_dim, hidden), np.zeros(hidden), w(hidden, hidden), np.zeros(hidden),
                 w(hidden, out_dim) * 0.1, np.zeros(out_dim)]
        return np.concatenate([p.ravel() for p in parts]).astype(np.float32)

    def unpack(self):
        """(W1, b1, W2, b2, W3, b3) as float64 numpy arrays."""
        p = self.params.double().cpu().numpy()
        i, h, o = self.in_dim, self.hidden, self.out_dim
        sizes = [(i, h), (h,), (h, h), (h,), (h, o), (o,)]
        out, k = [], 0
        for s in sizes:
            n = int(np.prod(s))
            out.append(p[k:k + n].reshape(s))
            k += n
        return out

    def forward(self, X):
        B = X.shape[0]
        H1 = torch.empty((B, self.hidden), dtype=torch.float32, device=self.device)
        H2 = torch.empty_like(H1)
        Y = torch.empty((B, self.out_dim), dtype=torch.float32, device=self.device)
        _lib.mlp_forward(self.net, X.contiguous(), H1, H2, Y, self.device)
        return H1, H2, Y

    def loss_and_grad(self, X, T, M):
        H1, H2, Y = self.forward(X)
        grad = torch.empty_like(self.params)
        loss = torch.empty(1, dtype=torch.float32, device=self.device)
        ws = torch.empty(_lib.mlp_workspace_bytes(self.in_dim, self.hidden, self.out_dim, X.shape[0]),
                         dtype=torch.uint8, device=self.device)
        _lib.mlp_backward(self.net, X, H1, H2, Y, T.contiguous(), M.contiguous(), grad, loss, ws,
                          self.device)
        return loss, grad

    def adam(self, grad, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        _lib.adam_step(self.params, grad, self.m, self.v, lr, beta1, beta2, eps, self.step,
                       self.device)

    def fit(self, X, T, M, epochs=3000, lr=3e-3, log=None, log_every=500):
        losses = []
        for e in range(epochs):
            loss, grad = self.loss_and_grad(X, T, M)
            self.adam(grad, lr=lr * (0.5 * (1 + math.cos(math.pi * e / epochs)) * 0.9 + 0.1))
            if e % log_every == 0 or e == epochs - 1:
                lv = float(loss.item())
                losses.append((e, lv))
                if log:
                    log(f"epoch {e} loss {lv:.5f}")
        return losses


def targets_from_times(times: np.ndarray):
    """log2(t / min t) per row; mask 0 where a time is missing / non-finite."""
    t = np.asarray(times, np.float64)
    ok = np.isfinite(t) & (t > 0)
    best = np.min(np.where(ok, t, np.inf), axis=1, keepdims=True)
    T = np.where(ok, np.log2(np.where(ok, t, 1.0) / best), 0.0)
    return T.astype(np.float32), ok.astype(np.float32)


class Selector:
    """A trained model for one operation and variant list."""

    def __init__(self, op: str, mlp: Mlp, mu: np.ndarray, sd: np.ndarray, meta: dict | None = None):
        self.op = op
        self.mlp = mlp
        self.variants = SPMV_VARIANTS if op == "spmv" else SPMM_VARIANTS
        if mlp.out_dim != len(self.variants):
            raise ValueError("model outputs do not match the variant list")
        self.mu = torch.as_tensor(np.asarray(mu, np.float32), device=mlp.device)
        self.sd = torch.as_tensor(np.asarray(sd, np.float32), device=mlp.device)
        self.meta = dict(meta or {})

    @classmethod
    def train(cls, op, features, times, device, epochs=3000, lr=3e-3, seed=0, log=None):
        X = np.asarray(features, np.float64)
        mu = X.mean(axis=0)
        sd = X.std(axis=0)
        sd[sd == 0] = 1.0
        Z = ((X - mu) / sd).astype(np.float32)
        T, M = targets_from_times(times)
        mlp = Mlp(T.shape[1], device, seed=seed)
        dev = mlp.device
        losses = mlp.fit(torch.from_numpy(Z).to(dev), torch.from_numpy(T).to(dev),
                         torch.from_numpy(M).to(dev), epochs=epochs, lr=lr, log=log)
        return cls(op, mlp, mu, sd, {"losses": losses, "n_train": int(X.shape[0])})

    def predict(self, features) -> np.ndarray:
        """Predicted log2 relative times for a batch of raw feature rows."""
        X = np.asarray(features, np.float64).reshape(-1, IN_DIM)
        Z = (X - self.mu.double().cpu().numpy()) / self.sd.double().cpu().numpy()
        _, _, Y = self.mlp.forward(torch.from_numpy(Z.astype(np.float32)).to(self.mlp.device))
        return Y.double().cpu().numpy()

    def covers(self, n_rows: int, n_cols: int, nnz: int) -> bool:
        """Whether the matrix lies inside the training set's size range
        (meta "n_max" / "nnz_max"; a model without them covers everything)."""
        n_max = self.meta.get("n_max")
        nnz_max = self.meta.get("nnz_max")
        if n_max is not None and max(n_rows, n_cols) > 1.1 * float(n_max):
            return False
        if nnz_max is not None and nnz > 1.1 * float(nnz_max):
            return False
        return True

    def select_on_device(self, feats_dev):
        """Variant index (device int32 tensor) for one feature vector on the device."""
        choice = torch.empty(1, dtype=torch.int32, device=self.mlp.device)
        _lib.mlp_select(self.mlp.net, feats_dev, self.mu, self.sd, choice, None, self.mlp.device)
        return choice

    def schedule_for(self, stats_dev, n_rows, n_cols, nnz):
        """Features + selection on the device, one host read; returns
        (Schedule, index)."""
        f = _lib.features(stats_dev, n_rows, n_cols, nnz, self.mlp.device)
        idx = int(self.select_on_device(f).item())
        return self.variants[idx], idx

    def to_json(self) -> dict:
        return {
            "op": self.op,
            "variants_version": VARIANTS_VERSION[self.op],
            "in_dim": self.mlp.in_dim, "hidden": self.mlp.hidden, "out_dim": self.mlp.out_dim,
            "params": self.mlp.params.cpu().numpy().astype(float).tolist(),
            "mu": self.mu.cpu().numpy().astype(float).tolist(),
            "sd": self.sd.cpu().numpy().astype(float).tolist(),
            "meta": self.meta,
        }

    def save(self, path: str):
        with open(path, "w") as fh:
            json.dump(self.to_json(), fh)

    @classmethod
    def load(cls, path: str, device) -> "Selector":
        with open(path) as fh:
            d = json.load(fh)
        if int(d["variants_version"]) != VARIANTS_VERSION.get(d["op"], -1):
            raise ValueError(f"{path}: variant list version mismatch")
        mlp = Mlp(int(d["out_dim"]), device, hidden=int(d["hidden"]), in_dim=int(d["in_dim"]),
                  params=np.asarray(d["params"], np.float32))
        return cls(d["op"], mlp, np.asarray(d["mu"]), np.asarray(d["sd"]), d.get("meta"))


_CACHE = {}


def model_path(op: str, value_tag: str = "f32") -> str:
    """Packaged model of op trained on timings with `value_tag` values
    ("f32", or "f64" for the SpMV model trained on fp64 timings)."""
    return os.path.join(default_db_dir(), f"{op}_{value_tag}_mlp.json")


def load_default(op: str, device, value_tag: str = "f32"):
    """The packaged model for op ("spmv" / "spmm64") and value tag, or None
    if absent or trained for another variant list (a warning says so;
    callers then use the fixed rule)."""
    key = (op, str(device), value_tag)
    if key not in _CACHE:
        path = model_path(op, value_tag)
        model = None
        if os.path.exists(path):
            try:
                model = Selector.load(path, device)
            except ValueError as exc:
                warnings.warn(f"schedule selector not used: {exc}", RuntimeWarning, stacklevel=2)
        _CACHE[key] = model
    return _CACHE[key]
