"""Collect autotuning timings, train the schedule selectors, evaluate them.

    python -m flagsparse_public_amd.train_selector --out DIR [--n-train 2000]
        [--n-test 200] [--max-n 2000000] [--resume]

Matrices: seeds 5000.. (training) and 7000.. (held out), five families
(``synthetic.SELECTOR_FAMILIES``) with n log-uniform in [1e4, 2e6], plus a
large-matrix supplement with n log-uniform in [2e6, 5e6] (seeds 9000..
training, 9500.. held out), fp32 values, generated on the GPU.  For every matrix all SpMV variants and all
64-column SpMM variants are timed (``autotune``); the selectors are trained
on the training rows with the library's HIP MLP kernels; on the held-out
matrices the on-device selection is compared with exhaustive autotuning
(regret) and with the fixed default schedule (speed-up).

Writes DIR/{spmv,spmm64}_db.npz, DIR/{spmv,spmm64}_f32_mlp.json and
DIR/eval.json; prints one progress line per 20 matrices.

    python -m flagsparse_public_amd.train_selector --out DIR --eval-only

scores the packaged models on the held-out timings in DIR again (no
collection, no training) and rewrites their reports in DIR/eval.json.

    python -m flagsparse_public_amd.train_selector --out DIR --value-dtype float64 --ops spmv

does the same for SpMV with fp64 values and x: the same matrices, timed with
fp64 values (DIR/spmv_f64_{train,test}_db.npz), train the model the library
uses for fp64 SpMV (DIR/spmv_f64_mlp.json, eval.json "spmv_f64").

    python -m flagsparse_public_amd.train_selector --out DIR --value-dtype float64 --eval-only

times every SpMV variant on the held-out matrices with fp64 values
(DIR/spmv_f64_test_db.npz, unless --resume finds it) and scores on them the
packaged fp32-trained SpMV model (eval.json "spmv_f64_held_out") and the
fp64-trained one when packaged (eval.json "spmv_f64"): the features depend
only on the sparsity structure, so the first question is whether the best
variant does too.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np
import torch

from . import _lib, synthetic
from .autotune import (SPMM_DEFAULT_INDEX, SPMM_VARIANTS, SPMV_DEFAULT_INDEX,
                       SPMV_FIXED_S8_INDEX, SPMV_VARIANTS, TimingDB, spmv_heuristic_index,
                       tune_spmm, tune_spmv)
from .selector import Selector
from .sparse_operations.spmm_csr import prepare_spmm_csr_opt
from .sparse_operations.spmv_csr import prepare_spmv_csr

TRAIN_SEED0 = 5000
TEST_SEED0 = 7000
LARGE_TRAIN_SEED0 = 9000
LARGE_TEST_SEED0 = 9500


def _log(msg):
    print(f"[train_selector] {msg}", flush=True)


def collect(seeds, device, dbs, max_n, reps, t0, label, save=None, save_every=100,
            deadline=None, value_dtype=torch.float32):
    """Time every variant of each operation in ``dbs`` on each matrix (values
    and x of ``value_dtype``); returns False when the deadline (time.time()
    value) stopped collection early."""
    dense_cache = {}
    for k, seed in enumerate(seeds):
        if deadline is not None and time.time() > deadline:
            _log(f"{label}: time budget reached after {k} matrices")
            return False
        family, n = synthetic.selector_matrix_spec(seed)
        n = min(n, max_n)
        indptr, cols, vals, params = synthetic.family_csr_torch(family, n, seed, device)
        vals = vals.to(value_dtype)
        nnz = int(vals.numel())
        pv = prepare_spmv_csr(vals, cols, indptr, (n, n), schedule="default")
        stats = _lib.analyze_device(pv.csr_struct(), device)
        feats = _lib.features(stats, n, n, nnz, device).double().cpu().numpy()
        g = torch.Generator(device=device)
        g.manual_seed(seed + 1)
        x = (torch.rand(n, generator=g, device=device) * 2 - 1).to(value_dtype)
        best = []
        if "spmv" in dbs:
            tv = tune_spmv(pv, x, reps=reps)
            dbs["spmv"].add(seed, family, feats, tv, n, nnz)
            best.append(f"spmv best {tv.min():.1f}us")
        if "spmm64" in dbs:
            if n not in dense_cache:
                dense_cache.clear()
                dense_cache[n] = torch.rand((n, 64), generator=g, device=device) * 2 - 1
            pm = prepare_spmm_csr_opt(vals, cols, indptr, (n, n), schedule="default")
            tm = tune_spmm(pm, dense_cache[n], reps=max(3, reps - 2))
            dbs["spmm64"].add(seed, family, feats, tm, n, nnz)
            best.append(f"spmm best {tm.min():.1f}us")
            del pm
        del pv, indptr, cols, vals, x
        if save is not None and (k + 1) % save_every == 0:
            save()
        if (k + 1) % 20 == 0 or k + 1 == len(seeds) or n >= 1_000_000:
            _log(f"{label} {k + 1}/{len(seeds)} matrices, {time.time() - t0:.0f}s "
                 f"(last: {family} n={n} nnz={nnz}, {', '.join(best)})")
    return True


def evaluate(sel: Selector, db: TimingDB, default_index: int, device, heuristic=None):
    """Regret against exhaustive autotuning and speed-up over the plain
    default-tile variant (and over the fixed heuristic rule when given: a
    function (n_cols, n_rows, nnz, local) -> variant index, ``local`` the
    family's column locality, synthetic.LOCAL_FAMILIES)."""
    _, fams, feats, times, sizes = db.arrays()
    choices = []
    for f in feats:
        fd = torch.as_tensor(f.astype(np.float32), device=device)
        choices.append(int(sel.select_on_device(fd).item()))
    choices = np.asarray(choices)
    rows = np.arange(len(choices))
    t_sel = times[rows, choices]
    t_best = times.min(axis=1)
    t_def = times[:, default_index]
    regret = t_sel / t_best
    speedup = t_def / t_sel
    out = {
        "n": int(len(choices)),
        "geomean_regret": float(np.exp(np.mean(np.log(regret)))),
        "frac_within_5pct_of_exhaustive": float(np.mean(regret <= 1.05)),
        "total_time_selected_over_exhaustive": float(t_sel.sum() / t_best.sum()),
        "geomean_speedup_vs_default": float(np.exp(np.mean(np.log(speedup)))),
        "total_time_default_over_selected": float(t_def.sum() / t_sel.sum()),
        "exhaustive_geomean_speedup_vs_default": float(np.exp(np.mean(np.log(t_def / t_best)))),
        "per_family": {},
    }
    if heuristic is not None:
        hidx = np.asarray([heuristic(int(n), int(n), int(z), str(f) in synthetic.LOCAL_FAMILIES)
                           for (n, z), f in zip(sizes, fams)])
        t_heu = times[rows, hidx]
        out["geomean_speedup_vs_heuristic"] = float(np.exp(np.mean(np.log(t_heu / t_sel))))
        out["heuristic_geomean_regret"] = float(np.exp(np.mean(np.log(t_heu / t_best))))
        t_s8 = times[:, SPMV_FIXED_S8_INDEX]
        out["geomean_speedup_vs_fixed_s8"] = float(np.exp(np.mean(np.log(t_s8 / t_sel))))
        out["fixed_s8_geomean_regret"] = float(np.exp(np.mean(np.log(t_s8 / t_best))))
    for fam in sorted(set(fams.tolist())):
        m = fams == fam
        out["per_family"][fam] = {
            "geomean_regret": float(np.exp(np.mean(np.log(regret[m])))),
            "geomean_speedup_vs_default": float(np.exp(np.mean(np.log(speedup[m])))),
        }
    return out, choices


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", required=True)
    ap.add_argument("--n-train", type=int, default=2000)
    ap.add_argument("--n-test", type=int, default=200)
    ap.add_argument("--n-large", type=int, default=100,
                    help="training matrices of the large supplement (n in [2e6, 5e6])")
    ap.add_argument("--n-large-test", type=int, default=10)
    ap.add_argument("--max-n", type=int, default=5_000_000)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--epochs", type=int, default=4000)
    ap.add_argument("--resume", action="store_true", help="reuse DIR/*_db.npz if present")
    ap.add_argument("--collect-only", action="store_true")
    ap.add_argument("--time-budget", type=float, default=0.0,
                    help="stop collecting after this many seconds (saved; resume later)")
    ap.add_argument("--ops", default="spmv,spmm64",
                    help="operations to collect and train (the others' files are kept)")
    ap.add_argument("--value-dtype", choices=["float32", "float64"], default="float32")
    ap.add_argument("--eval-only", action="store_true",
                    help="score the packaged models without training: on the held-out timings "
                         "in DIR (float32), or on held-out matrices timed in fp64 (float64)")
    args = ap.parse_args(argv)
    if args.eval_only:
        return eval_f64(args) if args.value_dtype == "float64" else eval_f32(args)
    ops = [o for o in args.ops.split(",") if o]
    for o in ops:
        if o not in ("spmv", "spmm64"):
            raise SystemExit(f"unknown op {o}")
    f64 = args.value_dtype == "float64"
    if f64 and ops != ["spmv"]:
        raise SystemExit("fp64 timings are collected for SpMV only: use --ops spmv")
    tag = "f64" if f64 else "f32"
    value_dtype = torch.float64 if f64 else torch.float32
    os.makedirs(args.out, exist_ok=True)
    device = torch.device("cuda", 0)
    name = torch.cuda.get_device_name(device)
    t0 = time.time()
    db_tag = "_f64" if f64 else ""
    paths = {s: {op: os.path.join(args.out, f"{op}{db_tag}_{s}_db.npz") for op in ops}
             for s in ("train", "test")}
    n_variants = {"spmv": len(SPMV_VARIANTS), "spmm64": len(SPMM_VARIANTS)}
    dbs = {}
    deadline = t0 + args.time_budget if args.time_budget > 0 else None
    complete = True
    # held-out matrices first, so a partial collection can already be judged
    splits = (("test", [(TEST_SEED0, args.n_test), (LARGE_TEST_SEED0, args.n_large_test)]),
              ("train", [(TRAIN_SEED0, args.n_train), (LARGE_TRAIN_SEED0, args.n_large)]))
    for split, ranges in splits:
        have = {}
        if args.resume and all(os.path.exists(p) for p in paths[split].values()):
            have = {op: TimingDB.load(p) for op, p in paths[split].items()}
        if not have:
            have = {op: TimingDB(op, n_variants[op], name) for op in ops}
        done = set(have[ops[0]].ids)
        seeds = [s for seed0, count in ranges for s in range(seed0, seed0 + count) if s not in done]
        def save(have=have, split=split):
            for op, db in have.items():
                db.save(paths[split][op])
        if seeds and complete:
            complete = collect(seeds, device, have, args.max_n, args.reps, t0, f"{split} {tag}",
                               save=save, deadline=deadline, value_dtype=value_dtype)
            save()
        dbs[split] = have
    if args.collect_only or not complete:
        _log(f"collected: test {len(dbs['test'][ops[0]])}, train {len(dbs['train'][ops[0]])}")
        return 0
    eval_path = os.path.join(args.out, "eval.json")
    report = {}
    if os.path.exists(eval_path):  # keep the other operations' reports
        with open(eval_path) as fh:
            report = json.load(fh)
    if not f64:
        report.update({"device": name, "n_train": len(dbs["train"][ops[0]]),
                       "n_test": len(dbs["test"][ops[0]])})
    defaults = {"spmv": SPMV_DEFAULT_INDEX, "spmm64": SPMM_DEFAULT_INDEX}
    for op in ops:
        default_index = defaults[op]
        _, _, Xtr, Ttr, _ = dbs["train"][op].arrays()
        sel = Selector.train(op, Xtr, Ttr, device, epochs=args.epochs,
                             log=lambda m, op=op: _log(f"{op}: {m}"))
        _, _, _, _, sizes = dbs["train"][op].arrays()
        sel.meta.update({"device": name, "default_index": default_index,
                         "n_max": int(sizes[:, 0].max()), "nnz_max": int(sizes[:, 1].max())})
        heur = (_heuristic_f64 if f64 else _heuristic_f32) if op == "spmv" else None
        ev, _ = evaluate(sel, dbs["test"][op], default_index, device, heur)
        tr, _ = evaluate(sel, dbs["train"][op], default_index, device, heur)
        report[op + db_tag] = {"held_out": ev, "train": {k: tr[k] for k in
                                                         ("geomean_regret", "geomean_speedup_vs_default")},
                               "n_train": len(dbs["train"][op]), "n_test": len(dbs["test"][op])}
        sel.meta["held_out"] = ev
        sel.meta["value_dtype"] = args.value_dtype
        sel.save(os.path.join(args.out, f"{op}_{tag}_mlp.json"))
        _log(f"{op}: held-out regret {ev['geomean_regret']:.3f}, "
             f"speed-up vs default {ev['geomean_speedup_vs_default']:.3f}")
    with open(eval_path, "w") as fh:
        json.dump(report, fh, indent=1)
    print(json.dumps(report), flush=True)
    return 0


def _heuristic_f32(n_cols, n_rows, nnz, local):
    return spmv_heuristic_index(n_cols, n_rows, nnz, local=local)


def _heuristic_f64(n_cols, n_rows, nnz, local):
    return spmv_heuristic_index(n_cols, n_rows, nnz, item=8, local=local)


def eval_f32(args):
    """Score the packaged models on the held-out timings already in DIR
    (DIR/{op}_test_db.npz) without collecting or training: rewrites each
    operation's "held_out" report in DIR/eval.json and the model's copy of
    it."""
    from .selector import load_default
    device = torch.device("cuda", 0)
    eval_path = os.path.join(args.out, "eval.json")
    report = {}
    if os.path.exists(eval_path):
        with open(eval_path) as fh:
            report = json.load(fh)
    defaults = {"spmv": SPMV_DEFAULT_INDEX, "spmm64": SPMM_DEFAULT_INDEX}
    for op in [o for o in args.ops.split(",") if o]:
        path = os.path.join(args.out, f"{op}_test_db.npz")
        if op not in defaults or not os.path.exists(path):
            raise SystemExit(f"no held-out timings for {op} in {args.out}")
        sel = load_default(op, device)
        if sel is None:
            raise SystemExit(f"no packaged {op} model to evaluate")
        heur = _heuristic_f32 if op == "spmv" else None
        ev, _ = evaluate(sel, TimingDB.load(path), defaults[op], device, heur)
        report.setdefault(op, {})["held_out"] = ev
        mpath = os.path.join(args.out, f"{op}_f32_mlp.json")
        if os.path.exists(mpath):  # the model's copy of its report; parameters untouched
            with open(mpath) as fh:
                d = json.load(fh)
            d.setdefault("meta", {})["held_out"] = ev
            with open(mpath, "w") as fh:
                json.dump(d, fh)
        _log(f"{op}: held-out regret {ev['geomean_regret']:.3f}, speed-up vs default "
             f"{ev['geomean_speedup_vs_default']:.3f}, vs fixed rule "
             f"{ev.get('geomean_speedup_vs_heuristic', 0):.3f}")
    with open(eval_path, "w") as fh:
        json.dump(report, fh, indent=1)
    return 0


def eval_f64(args):
    """Held-out matrices with fp64 values: time every SpMV variant (or reuse
    the timings with --resume), score the packaged fp32-trained model, the
    fp64-trained one when packaged, and the fp64 fixed rule on them."""
    from .selector import load_default
    os.makedirs(args.out, exist_ok=True)
    device = torch.device("cuda", 0)
    name = torch.cuda.get_device_name(device)
    t0 = time.time()
    path = os.path.join(args.out, "spmv_f64_test_db.npz")
    db = TimingDB.load(path) if (args.resume and os.path.exists(path)) else \
        TimingDB("spmv", len(SPMV_VARIANTS), name)
    done = set(db.ids)
    seeds = [s for seed0, count in ((TEST_SEED0, args.n_test), (LARGE_TEST_SEED0, args.n_large_test))
             for s in range(seed0, seed0 + count) if s not in done]
    deadline = t0 + args.time_budget if args.time_budget > 0 else None
    complete = collect(seeds, device, {"spmv": db}, args.max_n, args.reps, t0, "test f64",
                       save=lambda: db.save(path), deadline=deadline,
                       value_dtype=torch.float64)
    db.save(path)
    if not complete:
        _log(f"collected: {len(db)} fp64 held-out matrices")
        return 0
    sel = load_default("spmv", device)
    if sel is None:
        raise SystemExit("no packaged SpMV model to evaluate")
    ev, _ = evaluate(sel, db, SPMV_DEFAULT_INDEX, device, _heuristic_f64)
    eval_path = os.path.join(args.out, "eval.json")
    report = {}
    if os.path.exists(eval_path):
        with open(eval_path) as fh:
            report = json.load(fh)
    report["spmv_f64_held_out"] = ev
    sel64 = load_default("spmv", device, "f64")
    if sel64 is not None:
        ev64, _ = evaluate(sel64, db, SPMV_DEFAULT_INDEX, device, _heuristic_f64)
        report.setdefault("spmv_f64", {})["held_out"] = ev64
        _log(f"fp64-trained model: regret {ev64['geomean_regret']:.3f}, within 5% "
             f"{ev64['frac_within_5pct_of_exhaustive']:.3f}, vs fixed rule "
             f"{ev64.get('geomean_speedup_vs_heuristic', 0):.3f}")
    with open(eval_path, "w") as fh:
        json.dump(report, fh, indent=1)
    _log(f"fp64: regret {ev['geomean_regret']:.3f}, within 5% {ev['frac_within_5pct_of_exhaustive']:.3f}, "
         f"vs fixed rule {ev.get('geomean_speedup_vs_heuristic', 0):.3f}")
    print(json.dumps(ev), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
