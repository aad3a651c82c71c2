"""What the trained selector gains, checked on the CPU from committed data.

The timing databases (flagsparse_public_amd/selector_models/*_db.npz) hold the
autotuner's per-variant device times for held-out matrices (seeds 7000..) and
the training matrices (5000..); the models (*_f32_mlp.json, and
spmv_f64_mlp.json on the same matrices timed with fp64 values) were trained on
the training rows by the library's HIP MLP kernels.  Here the oracle's MLP forward
(oracle/csr_oracle.c) recomputes every held-out choice from the stored
features, and the test checks
  * the selection matches the one recorded on the device at evaluation time
    (same regret / speed-up as eval.json, to rounding);
  * regret against exhaustive autotuning (time of the chosen variant over the
    fastest) and the speed-up over the plain default-tile variant and over
    the fixed heuristic rule -- the gain of the network over the methods
    without it.
"""

import json
import os

import numpy as np
import pytest

from flagsparse_public_amd.autotune import (SPMM_DEFAULT_INDEX, SPMV_DEFAULT_INDEX, TimingDB,
                                            default_db_dir, spmv_heuristic_index)
from flagsparse_public_amd.synthetic import LOCAL_FAMILIES

DIR = default_db_dir()


def _have(*names):
    return all(os.path.exists(os.path.join(DIR, n)) for n in names)


def _unpack(params, i, h, o):
    p = np.asarray(params, np.float64)
    sizes = [(i, h), (h,), (h, h), (h,), (h, o), (o,)]
    out, k = [], 0
    for s in sizes:
        n = int(np.prod(s))
        out.append(p[k:k + n].reshape(s))
        k += n
    return out


def _choices(orc, model, feats):
    mu = np.asarray(model["mu"], np.float64)
    sd = np.asarray(model["sd"], np.float64)
    # the device standardises fp32 features with fp32 mu / sd
    Z = ((feats.astype(np.float32) - mu.astype(np.float32)) / sd.astype(np.float32))
    W = _unpack(model["params"], model["in_dim"], model["hidden"], model["out_dim"])
    _, _, Y = orc.mlp_forward(Z.astype(np.float64), *W)
    return np.argmin(Y, axis=1), Y


@pytest.mark.parametrize("op,default_index", [("spmv", SPMV_DEFAULT_INDEX),
                                              ("spmm64", SPMM_DEFAULT_INDEX)])
def test_selector_gain_on_held_out(orc, op, default_index):
    if not _have(f"{op}_test_db.npz", f"{op}_f32_mlp.json", "eval.json"):
        pytest.skip("no trained selector committed for " + op)
    db = TimingDB.load(os.path.join(DIR, f"{op}_test_db.npz"))
    with open(os.path.join(DIR, f"{op}_f32_mlp.json")) as fh:
        model = json.load(fh)
    with open(os.path.join(DIR, "eval.json")) as fh:
        ev = json.load(fh)[op]["held_out"]
    _, fams, feats, times, sizes = db.arrays()
    choice, Y = _choices(orc, model, feats)
    rows = np.arange(times.shape[0])
    t_sel = times[rows, choice]
    t_best = times.min(axis=1)
    regret = float(np.exp(np.mean(np.log(t_sel / t_best))))
    speedup = float(np.exp(np.mean(np.log(times[:, default_index] / t_sel))))
    # the CPU forward reproduces the device's choices (ties within fp32 rounding aside)
    assert abs(regret - ev["geomean_regret"]) <= 2e-3 * ev["geomean_regret"]
    assert abs(speedup - ev["geomean_speedup_vs_default"]) <= 2e-3 * speedup
    # the network gains over the plain default-tile variant and stays close to
    # exhaustive autotuning
    assert speedup > 1.0
    assert regret < float(np.exp(np.mean(np.log(times[:, default_index] / t_best))))
    if op == "spmv":
        # the fixed rule as the library applies it (column locality sampled per
        # matrix: the banded and block-diagonal families are local)
        heur = np.asarray([spmv_heuristic_index(int(n), int(n), int(z), local=f in LOCAL_FAMILIES)
                           for (n, z), f in zip(sizes, fams)])
        t_heu = times[rows, heur]
        gain = float(np.exp(np.mean(np.log(t_heu / t_sel))))
        assert abs(gain - ev["geomean_speedup_vs_heuristic"]) <= 2e-3 * gain
        assert gain >= 0.97


def _fp64_regret(orc, model_name):
    db = TimingDB.load(os.path.join(DIR, "spmv_f64_test_db.npz"))
    with open(os.path.join(DIR, model_name)) as fh:
        model = json.load(fh)
    _, fams, feats, times, sizes = db.arrays()
    choice, _ = _choices(orc, model, feats)
    rows = np.arange(times.shape[0])
    return times[rows, choice] / times.min(axis=1)


# (model, its report in eval.json, regret bound, gain over the fp64 fixed rule)
FP64_MODELS = [("spmv_f32_mlp.json", ("spmv_f64_held_out",), 1.06, 1.05),
               ("spmv_f64_mlp.json", ("spmv_f64", "held_out"), 1.02, 1.10)]


@pytest.mark.parametrize("model_name,report,max_regret,min_gain", FP64_MODELS)
def test_spmv_model_on_fp64_held_out(orc, model_name, report, max_regret, min_gain):
    """Held-out matrices timed with fp64 values (spmv_f64_test_db.npz),
    scored for the fp32-trained SpMV model (its features depend only on the
    sparsity structure) and for the model trained on the same training
    matrices timed with fp64 values, the one the library uses for fp64
    values: the choices, recomputed here, are those recorded on the device,
    close to exhaustive autotuning and faster than the fp64 fixed rule."""
    if not _have("spmv_f64_test_db.npz", model_name, "eval.json"):
        pytest.skip("no fp64 evaluation committed")
    db = TimingDB.load(os.path.join(DIR, "spmv_f64_test_db.npz"))
    with open(os.path.join(DIR, model_name)) as fh:
        model = json.load(fh)
    with open(os.path.join(DIR, "eval.json")) as fh:
        ev = json.load(fh)
    for k in report:
        ev = ev[k]
    _, fams, feats, times, sizes = db.arrays()
    choice, _ = _choices(orc, model, feats)
    rows = np.arange(times.shape[0])
    t_sel = times[rows, choice]
    t_best = times.min(axis=1)
    regret = float(np.exp(np.mean(np.log(t_sel / t_best))))
    assert abs(regret - ev["geomean_regret"]) <= 2e-3 * ev["geomean_regret"]
    assert regret < max_regret
    heur = np.asarray([spmv_heuristic_index(int(n), int(n), int(z), item=8,
                                            local=f in LOCAL_FAMILIES)
                       for (n, z), f in zip(sizes, fams)])
    t_heu = times[rows, heur]
    gain = float(np.exp(np.mean(np.log(t_heu / t_sel))))
    assert abs(gain - ev["geomean_speedup_vs_heuristic"]) <= 2e-3 * gain
    assert gain > min_gain


def test_fp64_model_beats_fp32_model_on_fp64_held_out(orc):
    """On the fp64 held-out timings the fp64-trained model is closer to
    exhaustive autotuning than the fp32-trained one: lower geometric-mean
    regret and more matrices within 5% of the fastest variant."""
    if not _have("spmv_f64_test_db.npz", "spmv_f64_mlp.json", "spmv_f32_mlp.json"):
        pytest.skip("no fp64 model committed")
    r32 = _fp64_regret(orc, "spmv_f32_mlp.json")
    r64 = _fp64_regret(orc, "spmv_f64_mlp.json")
    assert np.exp(np.mean(np.log(r64))) < np.exp(np.mean(np.log(r32)))
    assert np.mean(r64 <= 1.05) > np.mean(r32 <= 1.05)
