#!/usr/bin/env python3
"""Benchmark: CSR SpMV on the 4M-row power-law workload, one MI355X.

Workload: n = 4,000,000 square, row lengths ceil of a continuous power law
(alpha 2, max 1e5, ~48.6M entries), uniform sorted distinct columns, fp32
values and x U(-1, 1), int32 indices and row pointer; y = A x.  A "step" is one
SpMV over the whole matrix with inputs resident in HBM.

Printed: one JSON line with GB/s over the algorithmic (gather-counted) bytes
  (n+1)*4 + nnz*4 (col) + nnz*4 (val) + nnz*4 (x gathers) + n*4 (y),
the roofline of the dominant kernel against 8 TB/s, and a CPU baseline (the
C oracle, single thread, on the host cores of the same box).

The default schedule stores this matrix column-slice-major (x is 16 MB, four
times one XCD's L2): a step is the compact-row SpMV (entry tiles, XCD-pinned
block groups, plus its small carry fix-up) and the combine pass.  The
dominant kernel is the compact-row SpMV; its own algorithmic bytes are
  (NC+1)*4 (compact row pointers) + nnz*12 (col, val, x gathers) + NC*4 (partials)
for NC compact rows, and its average duration is timed here with HIP events
on the launch stream, in a loop of its own after the timed steps.

The same JSON line nests, under "workloads", the rest of the metric, each
timed with HIP events in this run and carrying its error against an fp64
reference (SciPy's CSR product in float64):
  * spmv_powerlaw_4m_f64: the same matrix in fp64;
  * spmm_banded_1m: C = A B on the 1M-row banded matrix (offsets -3..3 plus 8
    random columns per row), N = 32 / 64 / 128 fp32 and N = 64 fp64, with a
    compulsory-bytes roofline (A once, B once, C once) and rocSPARSE beside it;
  * batched_512: 512 uneven matrices, one SpMV + one 8-column SpMM each, as one
    graph-replayed persistent launch against the 1,024 separate launches;
  * selection: on the 4M-row SpMV and the N = 64 SpMM, the learned choice
    against exhaustive autotuning (every variant timed) and the plain
    default-tile kernel;
  * unprepared_calls: flagsparse_spmv_csr(data, indices, indptr, x, shape)
    repeated on the same tensors (the reference's default call pattern);
  * modes_stream_bound: the split-precision modes beside plain fp32 on a
    4M-row banded matrix, where value bytes are a third of the traffic.
--quick skips them.

    python bench.py [--gpus 1] [--steps 200] [--warmup 20] [--path opt|default] [--dtype f32|f64]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0
# L2 request rate of MI355X: 8 XCDs x 16 L2 channels x one request per clock
# at the 2.4 GHz peak engine clock
L2_REQ_PEAK = 8 * 16 * 2.4e9
METRIC = "SpMV/SpMM GB/s & %HBM peak; learned+batched vs plain at fixed rel err"
ROOT = os.path.dirname(os.path.abspath(__file__))


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(n_rows, nnz, s_val=4, s_idx=4, s_ptr=4, s_x=4, s_y=4):
    return (n_rows + 1) * s_ptr + nnz * (s_idx + s_val + s_x) + n_rows * s_y


def _profile_value(name, key):
    """A `key = value` line of a committed rocprofv3 summary under profiles/."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + " ") or line.startswith(key + "="):
            try:
                return float(line.split("=", 1)[1].strip())
            except ValueError:
                return None
    return None


def _traffic_from_profile(name="spmv_powerlaw_4m_pmc.txt"):
    """HBM bytes per launch measured by rocprofv3 (FETCH_SIZE + WRITE_SIZE),
    scaled by the copy-kernel calibration, from the committed summary."""
    return _profile_value(name, "traffic_bytes_per_launch")


def cpu_baseline(ip, cols, vals, x, nbytes, budget_s=10.0):
    """Time the single-threaded C oracle on the same matrix (a bounded number
    of whole-matrix passes, about budget_s of CPU work); returns (dict, y64)."""
    sys.path.insert(0, ROOT)
    from oracle import oracle  # the checker, used only on this leg
    import numpy as np

    oracle.build()
    ip64 = np.ascontiguousarray(ip, np.int64)
    c64 = np.ascontiguousarray(cols, np.int64)
    v64 = np.ascontiguousarray(vals, np.float64)
    x64 = np.ascontiguousarray(x, np.float64)
    y64 = oracle.spmv(ip64, c64, v64, x64)  # warm
    reps, t_total = 0, 0.0
    while t_total < budget_s and reps < 200:
        t0 = time.perf_counter()
        oracle.spmv(ip64, c64, v64, x64)
        t_total += time.perf_counter() - t0
        reps += 1
    sec = t_total / reps
    res = {"value": round(nbytes / sec / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
           "sample": f"{reps} whole-matrix passes of the C oracle (fp64 accumulate, "
                     f"sequential, 1 thread), {sec * 1e3:.1f} ms each; GB/s over the same "
                     f"algorithmic bytes as the GPU line",
           "host_cpus": os.cpu_count(), "library_cpu_paths": library_cpu_paths(ip, cols, vals, x,
                                                                               nbytes)}
    return res, y64


def library_cpu_paths(ip, cols, vals, x, nbytes, reps=5):
    """The CPU library routines a user would otherwise call on the same fp32
    matrix: SciPy's CSR matvec (serial) and torch's CPU CSR product (all
    host threads); median of `reps` after one warm-up, GB/s over the same
    algorithmic bytes.  Informational, beside the oracle's figure."""
    import statistics
    import numpy as np
    out = []

    def timed(fn):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return statistics.median(ts)

    n = ip.size - 1
    try:
        import scipy.sparse as sp
        A = sp.csr_matrix((vals, cols, ip), shape=(n, n))
        sec = timed(lambda: A @ x)
        out.append({"library": f"scipy csr_matrix @ x ({vals.dtype}, serial)",
                    "ms": round(sec * 1e3, 2),
                    "GB/s": round(nbytes / sec / 1e9, 3), "threads": 1})
        del A
    except Exception as exc:  # reported, not fatal: a baseline only
        out.append({"library": "scipy", "error": str(exc)})
    try:
        import torch
        threads = os.cpu_count() or 1
        torch.set_num_threads(threads)
        import warnings
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", UserWarning)  # "sparse CSR support is in beta"
            A = torch.sparse_csr_tensor(torch.from_numpy(np.asarray(ip, np.int64)),
                                        torch.from_numpy(np.asarray(cols, np.int64)),
                                        torch.from_numpy(vals), size=(n, n))
        xt = torch.from_numpy(x)
        sec = timed(lambda: torch.mv(A, xt))
        out.append({"library": f"torch CPU CSR mv ({vals.dtype})", "ms": round(sec * 1e3, 2),
                    "GB/s": round(nbytes / sec / 1e9, 3), "threads": torch.get_num_threads()})
        if vals.dtype == np.float32:
            # the reference's own correctness oracle: torch sparse CSR in the
            # promoted dtype (fp32 -> fp64, tests/test_spmv.py:302-385)
            with warnings.catch_warnings():
                warnings.simplefilter("ignore", UserWarning)
                A64 = torch.sparse_csr_tensor(A.crow_indices(), A.col_indices(),
                                              A.values().to(torch.float64), size=(n, n))
            x64 = xt.to(torch.float64)
            sec = timed(lambda: torch.mv(A64, x64))
            out.append({"library": "torch CPU CSR mv, promoted to float64 (the reference's "
                                   "oracle)", "ms": round(sec * 1e3, 2),
                        "GB/s": round(nbytes / sec / 1e9, 3), "threads": torch.get_num_threads()})
            del A64
        del A
    except Exception as exc:
        out.append({"library": "torch CPU", "error": str(exc)})
    return out


def measure_modes(d_vals, d_cols, d_ip, d_x, n, nnz, prepared, use_opt, steps):
    """The same workload in the other modes, each timed like the main line:
    the default (fp64-accumulating) path and the three split-precision value
    storages.  Bytes per mode use that mode's stored value size."""
    import torch
    from flagsparse_public_amd import flagsparse_spmv_csr
    from flagsparse_public_amd.precision import (FORMAT_BYTES, flagsparse_spmv_csr_split,
                                                 prepare_spmv_csr_split)
    dev = d_x.device
    stream = torch.cuda.current_stream(dev)

    def timed(fn):
        for _ in range(3):
            fn()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record(stream)
        for _ in range(steps):
            out = fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) * 1e3 / steps, out

    out = {}
    us, yv = timed(lambda: flagsparse_spmv_csr(x=d_x, prepared=prepared, use_opt=not use_opt))
    b = spmv_bytes(n, nnz)
    out["default_path" if use_opt else "opt_path"] = {
        "us": round(us, 2), "GB/s": round(b / us / 1e3, 1), "_y": yv}
    for fmt in ("bf16x2", "bf16", "f16_rowscale"):
        ps = prepare_spmv_csr_split(d_vals, d_cols, d_ip, (n, n), fmt)
        us, yv = timed(lambda: flagsparse_spmv_csr_split(d_x, ps))
        per_entry, per_row = FORMAT_BYTES[fmt]
        b = (n + 1) * 4 + nnz * (4 + per_entry + 4) + n * (4 + per_row)
        out[fmt] = {"us": round(us, 2), "GB/s": round(b / us / 1e3, 1),
                    "value_bytes_per_entry": per_entry, "_y": yv}
        del ps
        _log(f"mode {fmt}: {us:.1f} us/step")
    return out


def _event_time(fn, steps, stream):
    import torch
    for _ in range(3):
        fn()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record(stream)
    for _ in range(steps):
        fn()
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) * 1e3 / steps


def _pass_times(passes, steps, stream):
    """Average duration (us) of each pass of a sequence run `steps` times in
    a row, each pass bracketed by HIP events on `stream`: every pass is
    timed where it runs in a real step (after the pass before it)."""
    import torch
    for _ in range(3):
        for f in passes:
            f()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(len(passes) + 1)]
           for _ in range(steps)]
    torch.cuda.synchronize()
    for k in range(steps):
        evs[k][0].record(stream)
        for i, f in enumerate(passes):
            f()
            evs[k][i + 1].record(stream)
    torch.cuda.synchronize()
    return [sum(e[i].elapsed_time(e[i + 1]) for e in evs) * 1e3 / steps
            for i in range(len(passes))]


def vendor_spmv(d_vals, d_cols, d_ip, d_x, n, nbytes, steps):
    """rocSPARSE CSR SpMV (adaptive algorithm, analysis done once) on the
    same matrix and x: the GPU vendor baseline."""
    import torch
    from flagsparse_public_amd import vendor
    if not vendor.available():
        return {"library": "rocSPARSE", "unavailable": vendor.unavailable_reason()}
    y = torch.empty(n, dtype=d_x.dtype, device=d_x.device)
    v = vendor.VendorSpmv(d_vals, d_cols, d_ip, (n, n), d_x, y, "adaptive")
    us = _event_time(v.run, steps, torch.cuda.current_stream(d_x.device))
    _log(f"rocSPARSE adaptive: {us:.1f} us/step")
    return {"library": "rocSPARSE CSR SpMV, adaptive", "us": round(us, 2),
            "GB/s": round(nbytes / us / 1e3, 1)}


def bench_spmm(args):
    """C = A B on the 1M-row banded matrix (offsets -3..3 plus 8 random
    columns per row), B row-major with --n-dense columns: GB/s over
    (n+1)*4 + nnz*(4 + s + N*s) + n*N*s (every B row gathered)."""
    import numpy as np
    import torch
    sys.path.insert(0, ROOT)
    from flagsparse_public_amd import (flagsparse_spmm_csr_opt, prepare_spmm_csr_opt, synthetic,
                                       vendor)
    from flagsparse_public_amd.sparse_operations.spmm_csr import spmm_launch_schedule
    dev = torch.device("cuda", 0)
    npd = np.float32 if args.dtype == "f32" else np.float64
    ip, cols, vals = synthetic.spmm_banded_1m(dtype=npd)
    n, nnz, N = ip.size - 1, int(ip[-1]), args.n_dense
    s_v = np.dtype(npd).itemsize
    d = [torch.from_numpy(a).to(dev) for a in (vals, cols, ip.astype(np.int32))]
    B = torch.from_numpy(synthetic.dense_uniform(303, (n, N), npd)).to(dev)
    p = prepare_spmm_csr_opt(*d, (n, n))
    launch = spmm_launch_schedule(p, N)
    stream = torch.cuda.current_stream(dev)
    us = _event_time(lambda: flagsparse_spmm_csr_opt(B=B, prepared=p), args.steps, stream)
    nbytes = (n + 1) * 4 + nnz * (4 + s_v + N * s_v) + n * N * s_v
    compulsory = (n + 1) * 4 + nnz * (4 + s_v) + 2 * n * N * s_v
    gbs = nbytes / us / 1e3
    res = {"metric": METRIC, "value": round(gbs, 2), "unit": "GB/s", "n_gpus": 1,
           "steps": args.steps, "warmup": 3, "ms_per_step": round(us / 1e3, 5),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": args.dtype,
           "data": "synthetic",
           "config": {"workload": "spmm_banded_1m", "n_rows": n, "nnz": nnz, "n_dense": N,
                      "gflops": round(2.0 * nnz * N / us / 1e3, 2),
                      "compulsory_bytes": compulsory,
                      "schedule": repr(launch[0]), "col_tile": launch[1]},
           # roofline over the compulsory bytes (A, B and C once each): the
           # gather-counted rate above exceeds the HBM peak because B rows are
           # re-read from the caches, so it bounds nothing
           "roofline": {"bound": "hbm", "kernel": "spmm_csr_kernel",
                        "achieved": round(compulsory / us / 1e3, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(compulsory / us / 1e3 / HBM_PEAK_GBS, 4),
                        "traffic": (_traffic_from_profile("spmm_banded_1m_n64_pmc.txt")
                                    if N == 64 and args.dtype == "f32" else None),
                        "compulsory_bytes_per_launch": compulsory,
                        "gather_counted_bytes_per_launch": nbytes,
                        "avg_launch_us": round(us, 3)}}
    if not args.no_vendor and vendor.available():
        C = torch.empty((n, N), dtype=B.dtype, device=dev)
        v = vendor.VendorSpmm(*d, (n, n), B, C)
        vus = _event_time(v.run, args.steps, stream)
        res["vendor_baseline"] = {"library": "rocSPARSE CSR SpMM", "us": round(vus, 2),
                                  "GB/s": round(nbytes / vus / 1e3, 1)}
    print(json.dumps(res), flush=True)
    return 0


def bench_sddmm(args):
    """out[p] = <x[row p], y[col p]> over a 1M x 1M pattern with Poisson(16)
    row lengths (16.8M entries), x and y [1M, K] row-major, K = --n-dense.
    GB/s over the gather-counted bytes nnz*(4 row id + 4 col + s out) +
    2*nnz*K*s (both rows read per entry); compulsory bytes count x and y once.
    The rocSPARSE SDDMM on the same inputs is the vendor line."""
    import numpy as np
    import torch
    sys.path.insert(0, ROOT)
    from flagsparse_public_amd import flagsparse_sddmm_csr, prepare_sddmm_csr, synthetic, vendor
    dev = torch.device("cuda", 0)
    dt = torch.float32 if args.dtype == "f32" else torch.float64
    n, K = 1 << 20, args.n_dense
    ip, cols, _ = synthetic.poisson_csr(n, 16.0, 1)
    nnz = int(ip[-1])
    g = torch.Generator(device=dev)
    g.manual_seed(11)
    x = torch.rand((n, K), generator=g, device=dev, dtype=dt) * 2 - 1
    y = torch.rand((n, K), generator=g, device=dev, dtype=dt) * 2 - 1
    p = prepare_sddmm_csr(torch.from_numpy(cols).to(dev), torch.from_numpy(ip).to(dev), (n, n),
                          k_hint=K)
    out = torch.empty(nnz, dtype=dt, device=dev)
    stream = torch.cuda.current_stream(dev)
    us = _event_time(lambda: flagsparse_sddmm_csr(x=x, y=y, prepared=p, out=out), args.steps,
                     stream)
    s_v = x.element_size()
    nbytes = nnz * (4 + 4 + s_v) + 2 * nnz * K * s_v
    gbs = nbytes / us / 1e3
    res = {"metric": METRIC, "value": round(gbs, 2), "unit": "GB/s", "n_gpus": 1,
           "steps": args.steps, "warmup": 3, "ms_per_step": round(us / 1e3, 5),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": args.dtype,
           "data": "synthetic",
           "config": {"workload": "sddmm_poisson_1m", "n_rows": n, "nnz": nnz, "k_dim": K,
                      "compulsory_bytes": nnz * (4 + 4 + s_v) + 2 * n * K * s_v},
           "roofline": {"bound": "hbm", "kernel": "sddmm_kernel", "achieved": round(gbs, 2),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                        "traffic": None, "algorithmic_bytes_per_launch": nbytes,
                        "avg_launch_us": round(us, 3)}}
    if not args.no_vendor and vendor.available():
        vout = torch.empty(nnz, dtype=dt, device=dev)
        v = vendor.VendorSddmm(p.indices, p.indptr, p.shape, x, y, vout)
        vus = _event_time(v.run, args.steps, stream)
        res["vendor_baseline"] = {"library": "rocSPARSE SDDMM", "us": round(vus, 2),
                                  "GB/s": round(nbytes / vus / 1e3, 1)}
    print(json.dumps(res), flush=True)
    return 0


def bench_batched(args):
    """512 uneven CSR matrices, each with one SpMV and one 8-column SpMM, in
    one persistent launch replayed from a HIP graph; beside it the 1,024
    separate launches.  GB/s over the batch's algorithmic bytes."""
    import torch
    sys.path.insert(0, ROOT)
    from flagsparse_public_amd.autotune import time_launches
    from flagsparse_public_amd.batched import BatchedCsr
    from flagsparse_public_amd.batched_workload import uneven_batch
    from flagsparse_public_amd import prepare_spmm_csr_opt, prepare_spmv_csr
    from flagsparse_public_amd.sparse_operations.spmm_csr import _hip_spmm_csr_impl_opt_prepared
    from flagsparse_public_amd.sparse_operations.spmv_csr import _hip_spmv_csr_impl_opt_prepared
    dev = torch.device("cuda", 0)
    mats, meta = uneven_batch(512, 4, dev)
    bat = BatchedCsr(mats)
    graph = bat.capture()
    us_graph = time_launches(graph.replay, reps=max(args.steps // 20, 5), warmup=2)
    us_eager = time_launches(bat.run, reps=max(args.steps // 20, 5), warmup=2)
    pv = [prepare_spmv_csr(m["data"], m["indices"], m["indptr"], m["shape"]) for m in mats]
    pm = [prepare_spmm_csr_opt(m["data"], m["indices"], m["indptr"], m["shape"]) for m in mats]
    ys = [torch.empty(m["shape"][0], device=dev) for m in mats]
    Cs = [torch.empty((m["shape"][0], 8), device=dev) for m in mats]

    def separate():
        for m, a, b, y, C in zip(mats, pv, pm, ys, Cs):
            _hip_spmv_csr_impl_opt_prepared(a, m["x"], out=y)
            _hip_spmm_csr_impl_opt_prepared(b, m["B"], out=C)
    us_sep = time_launches(separate, reps=3, warmup=1)
    gbs = bat.bytes / us_graph / 1e3
    res = {"metric": METRIC, "value": round(gbs, 2), "unit": "GB/s", "n_gpus": 1,
           "steps": max(args.steps // 20, 5), "warmup": 2, "ms_per_step": round(us_graph / 1e3, 5),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": "batched_512", **{k: v for k, v in meta.items()
                                                   if isinstance(v, (int, float, str))},
                      "eager_ms": round(us_eager / 1e3, 4),
                      "separate_launches_ms": round(us_sep / 1e3, 4),
                      "speedup_vs_separate": round(us_sep / us_graph, 3)},
           "roofline": {"bound": "hbm", "kernel": "batched_kernel", "achieved": round(gbs, 2),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                        "traffic": None, "algorithmic_bytes_per_launch": bat.bytes,
                        "avg_launch_us": round(us_graph, 3)}}
    print(json.dumps(res), flush=True)
    return 0


# ------------------------------------------------------------------ sub-results
#
# Every sub-result is timed with HIP events on the launch stream in this run
# and reports its error against an fp64 reference: SciPy's CSR product on the
# matrix promoted to float64 (the reference's own correctness oracle is a CSR
# product in the promoted dtype).  Error measure: max over rows of
# |got - ref| / sum_j |a_ij x_j| (SpMM: per element, sum_j |a_ij b_jk|).


def _scipy_csr(ip, cols, vals, shape):
    import numpy as np
    import scipy.sparse as sp
    return sp.csr_matrix((np.asarray(vals, np.float64), np.asarray(cols, np.int64),
                          np.asarray(ip, np.int64)), shape=shape)


def _rel_err(got, ref, cond):
    import numpy as np
    d = np.abs(np.asarray(got, np.float64) - ref)
    return float((d / np.where(cond > 0, cond, 1.0)).max()) if d.size else 0.0


def _err_report(got, ref, cond, value_dtype):
    """Per-row error of a result against the fp64 oracle, three ways: over
    the condition sum Σ|a x| (the bar the tests use), over |y64| (rows with
    y64 != 0; U(-1,1) data gives rows whose y is near zero, hence the median
    beside the max), and the reference's allclose error ratio max |d| / (atol
    + rtol |y64|) with its tolerances for the dtype (<= 1 passes allclose)."""
    import numpy as np
    from flagsparse_public_amd.sparse_operations._common import _tolerance_for_dtype
    d = np.abs(np.asarray(got, np.float64) - ref)
    atol, rtol = _tolerance_for_dtype(value_dtype)
    nz = ref != 0
    vs_y = d[nz] / np.abs(ref[nz]) if nz.any() else np.zeros(1)
    return {"max_rel_err_vs_cond": _rel_err(got, ref, cond),
            "median_rel_err_vs_abs_y": float(np.median(vs_y)),
            "max_rel_err_vs_abs_y": float(vs_y.max()),
            "allclose_error_ratio": float((d / (atol + rtol * np.abs(ref))).max()) if d.size else 0.0,
            "allclose_tolerance": {"atol": atol, "rtol": rtol}}


def sub_spmv_f64(steps):
    """The 4M-row power-law SpMV in fp64 (library schedule, native path)."""
    import numpy as np
    import torch
    from flagsparse_public_amd import flagsparse_spmv_csr, prepare_spmv_csr, synthetic
    dev = torch.device("cuda", 0)
    ip, cols, vals = synthetic.spmv_powerlaw_4m(dtype=np.float64)
    n, nnz = ip.size - 1, int(ip[-1])
    x = synthetic.dense_uniform(202, n, np.float64)
    d = [torch.from_numpy(a).to(dev) for a in (vals, cols, ip.astype(np.int32))]
    dx = torch.from_numpy(x).to(dev)
    p = prepare_spmv_csr(*d, (n, n))
    stream = torch.cuda.current_stream(dev)
    us = _event_time(lambda: flagsparse_spmv_csr(x=dx, prepared=p, use_opt=True), steps, stream)
    y = flagsparse_spmv_csr(x=dx, prepared=p, use_opt=True).cpu().numpy()
    A = _scipy_csr(ip, cols, vals, (n, n))
    ref = A @ x
    cond = abs(A) @ np.abs(x)
    nbytes = spmv_bytes(n, nnz, s_val=8, s_x=8, s_y=8)
    _log(f"spmv fp64: {us:.1f} us")
    rec = {"us": round(us, 2), "GB/s": round(nbytes / us / 1e3, 1),
           "frac": round(nbytes / us / 1e3 / HBM_PEAK_GBS, 4), "bytes": nbytes,
           "schedule": repr(p.schedule), "chosen_by": p.schedule_record["chosen_by"],
           "max_rel_err_vs_fp64": _rel_err(y, ref, cond)}
    del p
    torch.cuda.empty_cache()
    # the packaged selector is trained on fp32 timings; the best of all
    # variants timed on this fp64 matrix, for comparison with the fixed rule
    pe = prepare_spmv_csr(*d, (n, n), schedule="exhaustive")
    us_e = _event_time(lambda: flagsparse_spmv_csr(x=dx, prepared=pe, use_opt=True), steps, stream)
    ye = flagsparse_spmv_csr(x=dx, prepared=pe, use_opt=True).cpu().numpy()
    rec["exhaustive"] = {"us": round(us_e, 2), "GB/s": round(nbytes / us_e / 1e3, 1),
                         "schedule": repr(pe.schedule),
                         "max_rel_err_vs_fp64": _rel_err(ye, ref, cond)}
    _log(f"spmv fp64 exhaustive: {us_e:.1f} us ({pe.schedule})")
    del pe
    torch.cuda.empty_cache()
    return rec


def sub_spmm_c3(steps, no_vendor=False):
    """C = A B on the 1M-row banded matrix for N = 32 / 64 / 128 (fp32) and
    N = 64 (fp64), each with the schedule the library chooses."""
    import numpy as np
    import torch
    from flagsparse_public_amd import (flagsparse_spmm_csr_opt, prepare_spmm_csr_opt, synthetic,
                                       vendor)
    from flagsparse_public_amd.sparse_operations.spmm_csr import spmm_launch_schedule
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    out = {}
    for N, dt in ((32, "f32"), (64, "f32"), (128, "f32"), (64, "f64")):
        npd = np.float32 if dt == "f32" else np.float64
        ip, cols, vals = synthetic.spmm_banded_1m(dtype=npd)
        n, nnz = ip.size - 1, int(ip[-1])
        s_v = np.dtype(npd).itemsize
        d = [torch.from_numpy(a).to(dev) for a in (vals, cols, ip.astype(np.int32))]
        Bh = synthetic.dense_uniform(303, (n, N), npd)
        B = torch.from_numpy(Bh).to(dev)
        p = prepare_spmm_csr_opt(*d, (n, n))
        launch = spmm_launch_schedule(p, N)
        us = _event_time(lambda: flagsparse_spmm_csr_opt(B=B, prepared=p), steps, stream)
        C = flagsparse_spmm_csr_opt(B=B, prepared=p).cpu().numpy()
        A = _scipy_csr(ip, cols, vals, (n, n))
        ref = A @ Bh.astype(np.float64)
        cond = abs(A) @ np.abs(Bh.astype(np.float64))
        gather = (n + 1) * 4 + nnz * (4 + s_v + N * s_v) + n * N * s_v
        compulsory = (n + 1) * 4 + nnz * (4 + s_v) + n * N * s_v + n * N * s_v
        rec = {"us": round(us, 2), "GB/s_gather_counted": round(gather / us / 1e3, 1),
               "GB/s_compulsory": round(compulsory / us / 1e3, 1),
               "frac_compulsory": round(compulsory / us / 1e3 / HBM_PEAK_GBS, 4),
               "gflops": round(2.0 * nnz * N / us / 1e3, 1),
               "schedule": repr(launch[0]), "col_tile": launch[1],
               "chosen_by": p.schedule_record["chosen_by"],
               "max_rel_err_vs_fp64": _rel_err(C, ref, cond)}
        if not no_vendor and vendor.available():
            Cv = torch.empty((n, N), dtype=B.dtype, device=dev)
            v = vendor.VendorSpmm(*d, (n, n), B, Cv)
            vus = _event_time(v.run, steps, stream)
            rec["rocsparse_us"] = round(vus, 2)
        out[f"n{N}_{dt}"] = rec
        _log(f"spmm N={N} {dt}: {us:.1f} us")
        del p, B, d, A, ref, cond, C
        torch.cuda.empty_cache()
    return out


def sub_batched(steps):
    """512 uneven matrices (one SpMV, one 8-column SpMM each): graph replay of
    the persistent launch, the same eagerly, and the 1,024 separate launches."""
    import numpy as np
    import torch
    from flagsparse_public_amd.autotune import time_launches
    from flagsparse_public_amd.batched import BatchedCsr
    from flagsparse_public_amd.batched_workload import uneven_batch
    from flagsparse_public_amd import prepare_spmm_csr_opt, prepare_spmv_csr
    from flagsparse_public_amd.sparse_operations.spmm_csr import _hip_spmm_csr_impl_opt_prepared
    from flagsparse_public_amd.sparse_operations.spmv_csr import _hip_spmv_csr_impl_opt_prepared
    dev = torch.device("cuda", 0)
    mats, meta = uneven_batch(512, 4, dev)
    bat = BatchedCsr(mats)
    graph = bat.capture()
    reps = max(steps // 20, 5)
    us_graph = time_launches(graph.replay, reps=reps, warmup=2)
    us_eager = time_launches(bat.run, reps=reps, warmup=2)
    pv = [prepare_spmv_csr(m["data"], m["indices"], m["indptr"], m["shape"]) for m in mats]
    pm = [prepare_spmm_csr_opt(m["data"], m["indices"], m["indptr"], m["shape"]) for m in mats]
    ys = [torch.empty(m["shape"][0], device=dev) for m in mats]
    Cs = [torch.empty((m["shape"][0], 8), device=dev) for m in mats]

    def separate():
        for m, a, b, y, C in zip(mats, pv, pm, ys, Cs):
            _hip_spmv_csr_impl_opt_prepared(a, m["x"], out=y)
            _hip_spmm_csr_impl_opt_prepared(b, m["B"], out=C)
    us_sep = time_launches(separate, reps=3, warmup=1)
    graph.replay()
    torch.cuda.synchronize()
    err = 0.0
    for i in (0, 97, 255, 511):
        m = mats[i]
        ip = m["indptr"].cpu().numpy()
        A = _scipy_csr(ip, m["indices"].cpu().numpy(), m["data"].cpu().numpy(), m["shape"])
        x = m["x"].double().cpu().numpy()
        Bh = m["B"].double().cpu().numpy()
        err = max(err, _rel_err(bat.y[i].cpu().numpy(), A @ x, abs(A) @ np.abs(x)),
                  _rel_err(bat.C[i].cpu().numpy(), A @ Bh, abs(A) @ np.abs(Bh)))
    _log(f"batched: graph {us_graph:.0f} us, separate {us_sep:.0f} us")
    return {"matrices": meta["count"], "rows": meta["rows"], "nnz": meta["nnz"],
            "graph_replay_ms": round(us_graph / 1e3, 4), "eager_ms": round(us_eager / 1e3, 4),
            "separate_launches_ms": round(us_sep / 1e3, 4),
            "speedup_vs_separate": round(us_sep / us_graph, 3),
            "GB/s": round(bat.bytes / us_graph / 1e3, 1), "bytes": bat.bytes,
            "max_rel_err_vs_fp64": err}


def sub_selection(d_spmv, spmv_prepared, steps):
    """Learned choice vs exhaustive autotuning vs the plain default-tile
    kernel, on the 4M-row SpMV (16 variants) and the N = 64 banded SpMM
    (8 variants).  Every variant is timed the way the autotuner times it."""
    import numpy as np
    import torch
    from flagsparse_public_amd import prepare_spmm_csr_opt, synthetic
    from flagsparse_public_amd.autotune import (SPMM_DEFAULT_INDEX, SPMM_VARIANTS,
                                                SPMV_DEFAULT_INDEX, SPMV_VARIANTS,
                                                spmv_heuristic_index, tune_spmm, tune_spmv)
    from flagsparse_public_amd.rowsplit import columns_local
    dev = torch.device("cuda", 0)
    out = {}
    vals, cols, ip, x = d_spmv
    t = tune_spmv(spmv_prepared, x, reps=5)
    how = spmv_prepared.schedule_record["chosen_by"]
    li = int(how.split(":")[1]) if how.startswith("learned:") else None
    n = spmv_prepared.n_rows
    local = columns_local(ip, cols, n, n, vals.element_size())
    out["spmv_powerlaw_4m"] = _selection_record(t, li, SPMV_DEFAULT_INDEX,
                                                spmv_heuristic_index(n, n, int(vals.numel()),
                                                                     local=local),
                                                how)
    ipb, colsb, valsb = synthetic.spmm_banded_1m()
    nb = ipb.size - 1
    d = [torch.from_numpy(a).to(dev) for a in (valsb, colsb, ipb.astype(np.int32))]
    B = torch.from_numpy(synthetic.dense_uniform(303, (nb, 64))).to(dev)
    pm = prepare_spmm_csr_opt(*d, (nb, nb))
    tm = tune_spmm(pm, B, reps=5)
    howm = pm.schedule_record["chosen_by"]
    lm = int(howm.split(":")[1]) if howm.startswith("learned:") else None
    out["spmm_banded_1m_n64"] = _selection_record(tm, lm, SPMM_DEFAULT_INDEX, None, howm)
    _log(f"selection: spmv {out['spmv_powerlaw_4m']}, spmm {out['spmm_banded_1m_n64']}")
    return out


def sub_modes_banded(steps):
    """The split-precision modes on a stream-bound SpMV: a 4M-row band of 13
    diagonals (52M entries; every x gather hits L2), so the value bytes are a
    third of the traffic.  Each mode beside the plain fp32 kernel (library
    schedules), with its error against the fp64 reference."""
    import numpy as np
    import torch
    from flagsparse_public_amd import flagsparse_spmv_csr, prepare_spmv_csr, synthetic
    from flagsparse_public_amd.precision import (FORMAT_BYTES, flagsparse_spmv_csr_split,
                                                 prepare_spmv_csr_split)
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    ip, cols, vals = synthetic.banded_csr(4_000_000, 13, 6)
    n, nnz = ip.size - 1, int(ip[-1])
    x = synthetic.dense_uniform(14, n)
    d = [torch.from_numpy(a).to(dev) for a in (vals, cols, ip.astype(np.int32))]
    dx = torch.from_numpy(x).to(dev)
    A = _scipy_csr(ip, cols, vals, (n, n))
    ref, cond = A @ x.astype(np.float64), abs(A) @ np.abs(x.astype(np.float64))
    p = prepare_spmv_csr(*d, (n, n))
    us = _event_time(lambda: flagsparse_spmv_csr(x=dx, prepared=p, use_opt=True), steps, stream)
    y = flagsparse_spmv_csr(x=dx, prepared=p, use_opt=True).cpu().numpy()
    b = spmv_bytes(n, nnz)
    out = {"matrix": f"banded, 13 diagonals, n={n}, nnz={nnz}",
           "fp32": {"us": round(us, 2), "GB/s": round(b / us / 1e3, 1),
                    "schedule": repr(p.schedule), "max_rel_err_vs_fp64": _rel_err(y, ref, cond)}}
    for fmt in ("bf16x2", "bf16", "f16_rowscale"):
        ps = prepare_spmv_csr_split(*d, (n, n), fmt)
        us = _event_time(lambda: flagsparse_spmv_csr_split(dx, ps), steps, stream)
        ys = flagsparse_spmv_csr_split(dx, ps).cpu().numpy()
        per_entry, per_row = FORMAT_BYTES[fmt]
        bm = (n + 1) * 4 + nnz * (4 + per_entry + 4) + n * (4 + per_row)
        out[fmt] = {"us": round(us, 2), "GB/s": round(bm / us / 1e3, 1),
                    "speedup_vs_fp32": round(out["fp32"]["us"] / us, 3),
                    "max_rel_err_vs_fp64": _rel_err(ys, ref, cond)}
        del ps
    _log(f"modes on the banded matrix: {out}")
    return out


def sub_unprepared(d_spmv, prepared_us, reps=20):
    """Unprepared calls flagsparse_spmv_csr(data, indices, indptr, x, shape)
    -- the reference's default call pattern -- on the 4M-row matrix and on
    the 16K-row one: wall time per call (each call runs the validation pass
    and its host read, as the reference's does), beside the prepared step and
    the validation pass alone.  The first call prepares; later calls on the
    same tensors reuse it.  Timed with use_opt=True (the opt path, as the
    prepared step), and again with the reference's default use_opt=False
    (the default path: fp32 summed in fp64)."""
    import numpy as np
    import torch
    from flagsparse_public_amd import _lib, flagsparse_spmv_csr, prepare_spmv_csr, synthetic
    from flagsparse_public_amd.sparse_operations.spmv_csr import (_UNPREPARED,
                                                                  _prepare_spmv_csr_matrix)
    dev = torch.device("cuda", 0)
    out = {}

    def wall(fn, k):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e6 / k

    cases = [("spmv_powerlaw_4m", d_spmv, prepared_us)]
    ip, cols, vals = synthetic.spmv_16k()
    n1 = ip.size - 1
    d1 = tuple(torch.from_numpy(a).to(dev) for a in (vals, cols, ip))
    x1 = torch.from_numpy(synthetic.dense_uniform(11, n1)).to(dev)
    p1 = prepare_spmv_csr(*d1, (n1, n1))
    us1 = _event_time(lambda: flagsparse_spmv_csr(x=x1, prepared=p1, use_opt=True), 100,
                      torch.cuda.current_stream(dev))
    cases.append(("spmv_16k", d1 + (x1,), us1))
    for name, (vals_d, cols_d, ip_d, x_d), step_us in cases:
        n = int(ip_d.numel()) - 1
        _UNPREPARED.clear()
        call = lambda: flagsparse_spmv_csr(vals_d, cols_d, ip_d, x_d, (n, n), use_opt=True)
        first = wall(call, 1)
        again = wall(call, reps)
        call_default = lambda: flagsparse_spmv_csr(vals_d, cols_d, ip_d, x_d, (n, n))
        call_default()  # the default path's own layout, built once
        again_default = wall(call_default, reps)
        valid = wall(lambda: _prepare_spmv_csr_matrix(vals_d, cols_d, ip_d, (n, n),
                                                      _lengths=False), reps)
        csr = _lib.make_csr(vals_d, cols_d, ip_d, n, n)
        kern = _event_time(lambda: _lib.analyze_device(csr, dev), reps,
                           torch.cuda.current_stream(dev))
        out[name] = {"first_call_us": round(first, 1), "repeat_call_us": round(again, 1),
                     "repeat_call_default_path_us": round(again_default, 1),
                     "prepared_step_us": round(step_us, 2), "validation_pass_us": round(valid, 1),
                     "validation_kernel_us": round(kern, 1),
                     "repeat_over_step_plus_validation": round(again / (step_us + valid), 3)}
    _UNPREPARED.clear()
    _log(f"unprepared calls: {out}")
    return out


def _selection_record(times, learned, default, rule, how):
    best = int(times.argmin())
    rec = {"variants_timed": int(times.size), "exhaustive_best": best,
           "exhaustive_us": round(float(times[best]), 2),
           "variant_us": [round(float(t), 2) for t in times],
           "default_tile_us": round(float(times[default]), 2), "chosen_by": how}
    if learned is not None:
        rec["learned"] = learned
        rec["learned_us"] = round(float(times[learned]), 2)
        rec["learned_over_exhaustive"] = round(float(times[learned] / times[best]), 4)
        rec["speedup_learned_vs_default_tile"] = round(float(times[default] / times[learned]), 3)
    if rule is not None:
        rec["fixed_rule_us"] = round(float(times[rule]), 2)
    return rec


def device_stream_rates(mib=1024, reps=20):
    """What a plain streaming kernel reaches on this GPU: a device copy of
    `mib` MiB of fp32 (bytes read + written), averaged over `reps` launches
    timed with HIP events."""
    import torch
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    n = mib * (1 << 20) // 4
    src = torch.ones(n, device=dev)
    dst = torch.empty_like(src)
    copy_us = _event_time(lambda: dst.copy_(src), reps, stream)
    del src, dst
    torch.cuda.empty_cache()
    return {"copy_GBps": round(2 * n * 4 / copy_us / 1e3, 1), "copy_bytes": 2 * n * 4,
            "reps": reps}


def calibrate_copy(mib, reps):
    """Plain device-to-device copies of a known size: the reference point that
    turns rocprofv3 FETCH_SIZE / WRITE_SIZE readings into bytes."""
    import torch
    dev = torch.device("cuda", 0)
    n = mib * (1 << 20) // 4
    src = torch.rand(n, device=dev)
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "reps": reps}),
          flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--path", choices=["opt", "default"], default="opt")
    ap.add_argument("--n", type=int, default=4_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-modes", action="store_true",
                    help="skip the split-precision and default-path lines")
    ap.add_argument("--no-vendor", action="store_true", help="skip the rocSPARSE line")
    ap.add_argument("--quick", action="store_true",
                    help="only the headline SpMV line (no nested workloads)")
    ap.add_argument("--schedule", choices=["learned", "sliced"], default="learned",
                    help="the headline SpMV's schedule: the library's choice (default) or entry "
                         "tiles of 2048 over the column-sliced layout (2 MB of x per slice)")
    ap.add_argument("--workload", default="spmv_powerlaw_4m",
                    choices=["spmv_powerlaw_4m", "spmm_banded_1m", "batched_512",
                             "sddmm_poisson_1m"],
                    help="the SpMV headline workload (default), the banded SpMM, the batch "
                         "or the SDDMM")
    ap.add_argument("--n-dense", type=int, default=64,
                    help="spmm_banded_1m: columns of B; sddmm_poisson_1m: K")
    ap.add_argument("--dtype", choices=["f32", "f64"], default="f32",
                    help="value type of the workload (split-precision modes: fp32 only)")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="only run device copies of this many MiB (counter calibration)")
    args = ap.parse_args(argv)
    if args.calibrate_copy_mb:
        return calibrate_copy(args.calibrate_copy_mb, max(args.steps, 1))
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(the SpMV row-block sharding is described, not run)", file=sys.stderr)
        return 2
    if args.workload == "spmm_banded_1m":
        return bench_spmm(args)
    if args.workload == "batched_512":
        return bench_batched(args)
    if args.workload == "sddmm_poisson_1m":
        return bench_sddmm(args)

    import numpy as np
    import torch

    sys.path.insert(0, ROOT)
    from flagsparse_public_amd import flagsparse_spmv_csr, prepare_spmv_csr, synthetic

    if not torch.cuda.is_available():
        print("no HIP device", file=sys.stderr)
        return 1
    dev = torch.device("cuda", 0)
    t0 = time.time()
    npd = np.float32 if args.dtype == "f32" else np.float64
    s_v = np.dtype(npd).itemsize
    ip, cols, vals = synthetic.spmv_powerlaw_4m(n=args.n, dtype=npd)
    n = ip.size - 1
    nnz = int(ip[-1])
    x = synthetic.dense_uniform(202, n, npd)
    _log(f"generated n={n} nnz={nnz} max_row={int(np.diff(ip).max())} in {time.time() - t0:.1f}s")
    d_ip = torch.from_numpy(ip.astype(np.int32)).to(dev)
    d_cols = torch.from_numpy(cols).to(dev)
    d_vals = torch.from_numpy(vals).to(dev)
    d_x = torch.from_numpy(x).to(dev)
    from flagsparse_public_amd.rowsplit import L2_SLICE_BYTES, Schedule
    sched = (None if args.schedule == "learned" else
             Schedule(tile_nnz=2048, tile_mode=1, x_slice_cols=L2_SLICE_BYTES // s_v))
    prepared = prepare_spmv_csr(d_vals, d_cols, d_ip, (n, n), schedule=sched)
    use_opt = args.path == "opt"
    y = torch.empty(n, dtype=d_vals.dtype, device=dev)

    def step():
        return flagsparse_spmv_csr(x=d_x, prepared=prepared, use_opt=use_opt)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        y = step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - w0) * 1e3
    ev_ms = ev0.elapsed_time(ev1)
    per_launch_ms = ev_ms / args.steps
    nbytes = spmv_bytes(n, nnz, s_val=s_v, s_x=s_v, s_y=s_v)
    gbs = nbytes / (per_launch_ms * 1e-3) / 1e9
    _log(f"{args.steps} steps: events {ev_ms:.3f} ms ({per_launch_ms * 1e3:.2f} us/launch), "
         f"wall {wall_ms:.3f} ms, {gbs:.1f} GB/s")
    sch = prepared.schedule
    # the dominant kernel alone: the gather pass of the x-window layout, the
    # compact-row SpMV of the sliced layout, or the whole SpMV otherwise
    from flagsparse_public_amd import _lib
    from flagsparse_public_amd.sparse_operations.spmv_csr import xwin_plan_key
    acc = _lib.FS_ACC_NATIVE if use_opt else _lib.FS_ACC_WIDE
    sliced = prepared._plans.get(("sliced", sch.key(), acc))
    xw = prepared._plans.get(xwin_plan_key(prepared, sch, acc)[0])
    kernels = None
    if xw is not None:
        pst = xw.plan(acc).struct(1)
        yk = torch.empty(n, dtype=d_vals.dtype, device=dev)
        n_batches = -(-nnz // 64)
        run_table = xw.n_runs * 4 + n_batches * 12
        # gather pass: wcol (2 B) read and xg written per entry, x read once
        # (+ the run table for a tile-major xg)
        g_bytes = nnz * (2 + s_v) + n * s_v + (run_table if xw.order == "tile" else 0)
        # pass 2: row pointers, values, x_perm (2 B) and xg per entry, y
        # (+ the run table and tile_run for a window-major xg)
        p2_bytes = ((n + 1) * 4 + nnz * (2 * s_v + 2) + n * s_v
                    + (run_table + (xw.n_tiles + 1) * 4 if xw.order == "window" else 0))
        g_us, p2_us = _pass_times([lambda: _lib.xwin_gather(xw.layout, d_x, xw.xg, dev),
                                   lambda: _lib.spmv(xw.csr, xw.xg, yk, 1.0, 0.0, pst, dev)],
                                  args.steps, stream)
        g_name = "%s (x windows of %d columns in LDS, %s-major xg)" % (
            "xwin_gather_wm_kernel" if xw.order == "window" else "xwin_gather_kernel",
            xw.window, xw.order)
        p2_name = "spmv_entry_kernel gathered-x form (%s-major xg) + carry fix-up" % xw.order
        kernels = [{"kernel": g_name, "us": round(g_us, 3), "algorithmic_bytes": g_bytes},
                   {"kernel": p2_name, "us": round(p2_us, 3), "algorithmic_bytes": p2_bytes}]
        for k in kernels:
            k["GB/s"] = round(k["algorithmic_bytes"] / k["us"] / 1e3, 1)
            k["frac"] = round(k["GB/s"] / HBM_PEAK_GBS, 4)
        # the dominant kernel is the longer pass, timed inside the sequence
        kernel = None
        kernel_us, kbytes, kname = ((p2_us, p2_bytes, p2_name) if p2_us >= g_us
                                    else (g_us, g_bytes, g_name))
        nc = None
    elif sliced is not None:
        kplan = sliced.plan.struct(1)

        def kernel():
            _lib.spmv(sliced.csr, d_x, None if sliced.sums else sliced.z, 1.0, 0.0, kplan, dev)
        nc = sliced.n_compact
        kbytes = (nc + 1) * 4 + nnz * (4 + 2 * s_v) + nc * sliced.z.element_size()
        kname = "spmv_entry_kernel (compact rows) + carry fix-up"
    else:
        kernel = step
        nc = None
        kbytes = nbytes
        kname = "spmv_entry_kernel + carry fix-up"
    if kernel is not None:
        kernel_us = _event_time(kernel, args.steps, stream)
    kgbs = kbytes / (kernel_us * 1e-6) / 1e9
    _log(f"dominant kernel: {kernel_us:.2f} us/launch, {kgbs:.1f} GB/s over {kbytes} bytes")
    if kernels is not None:
        _log(f"kernels: {kernels}")
    result = {
        "metric": METRIC,
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {
            "workload": "spmv_powerlaw_4m",
            "n_rows": n, "nnz": nnz, "max_row_nnz": int(np.diff(ip).max()),
            "index_dtype": "int32", "path": args.path,
            "schedule": {"tile_mode": "entries" if sch.tile_mode == 1 else "rows",
                         "tile_nnz": sch.tile_nnz, "x_slice_cols": sch.x_slice_cols,
                         "n_slices": sliced.n_slices if sliced is not None else 1,
                         "compact_rows": nc,
                         "x_window": xw.window if xw is not None else 0,
                         "x_windows": xw.n_windows if xw is not None else 0,
                         "xg_order": xw.order if xw is not None else None,
                         "chosen_by": prepared.schedule_record["chosen_by"]},
            "gflops": round(2.0 * nnz / (per_launch_ms * 1e-3) / 1e9, 2),
            "compulsory_bytes": (n + 1) * 4 + nnz * (4 + s_v) + n * s_v + n * s_v,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kname,
            "achieved": round(kgbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(kgbs / HBM_PEAK_GBS, 4),
            "traffic": (_traffic_from_profile("spmv_powerlaw_4m_xwin_pmc.txt" if xw is not None
                                              else "spmv_powerlaw_4m_pmc.txt")
                        if args.dtype == "f32" else None),
            "algorithmic_bytes_per_launch": kbytes,
            "avg_launch_us": round(kernel_us, 3),
            "step_us": round(per_launch_ms * 1e3, 3),
            "step_frac": round(gbs / HBM_PEAK_GBS, 4),
        },
    }
    if kernels is not None:
        result["roofline"]["kernels"] = kernels
    # the rates plain streaming kernels reach on this GPU, beside the
    # published peak: how close the dominant kernel is to what is reachable
    streams = device_stream_rates()
    streams["kernel_frac_of_copy"] = round(kgbs / streams["copy_GBps"], 4)
    result["roofline"]["device_stream"] = streams
    # the second bound of a gather-heavy SpMV: L2 requests per launch (TCC_REQ,
    # rocprofv3, committed summary) over the launch time measured here
    reqs = (_profile_value("spmv_powerlaw_4m_xwin_pmc.txt" if xw is not None
                           else "spmv_powerlaw_4m_l2_counters.txt", "l2_requests_per_launch")
            if args.dtype == "f32" else None)
    if reqs:
        rate = reqs / (kernel_us * 1e-6)
        result["roofline_l2"] = {"bound": "l2_requests", "kernel": kname,
                                 "achieved": round(rate / 1e9, 2), "peak": L2_REQ_PEAK / 1e9,
                                 "unit": "G requests/s", "frac": round(rate / L2_REQ_PEAK, 4),
                                 "requests_per_launch": int(reqs)}
    if not args.no_vendor:
        result["vendor_baseline"] = vendor_spmv(d_vals, d_cols, d_ip, d_x, n, nbytes, args.steps)
    modes = None
    if not args.no_modes and args.dtype == "f32":  # split storage is an fp32 format
        modes = measure_modes(d_vals, d_cols, d_ip, d_x, n, nnz, prepared, use_opt,
                              max(args.steps // 4, 10))
    if not args.no_cpu:
        cb, y64 = cpu_baseline(ip, cols, vals, x, nbytes)
        result["cpu_baseline"] = cb
        from oracle import oracle
        cond = oracle.spmv_abs(ip, cols, vals, x)

        def rel_err(yd):
            yh = yd.double().cpu().numpy()
            return float((np.abs(yh - y64) / np.where(cond > 0, cond, 1.0)).max())
        result["config"]["max_rel_err_vs_fp64"] = rel_err(y)
        result["config"]["error_vs_fp64"] = _err_report(y.double().cpu().numpy(), y64, cond,
                                                        d_vals.dtype)
        if modes is not None:
            for m in modes.values():
                ym = m.pop("_y")
                m["max_rel_err_vs_fp64"] = rel_err(ym)
                m["error_vs_fp64"] = _err_report(ym.double().cpu().numpy(), y64, cond,
                                                 torch.float32)
    if modes is not None:
        for m in modes.values():
            m.pop("_y", None)
        result["config"]["modes"] = modes
    if not args.quick and args.dtype == "f32" and args.n == 4_000_000:
        sub_steps = max(min(args.steps, 100), 10)
        work = {}
        t_sub = time.time()
        work["selection"] = sub_selection((d_vals, d_cols, d_ip, d_x), prepared, sub_steps)
        prepared._plans.clear()  # the variants' layouts
        torch.cuda.empty_cache()
        work["unprepared_calls"] = sub_unprepared((d_vals, d_cols, d_ip, d_x),
                                                  per_launch_ms * 1e3)
        torch.cuda.empty_cache()
        work["modes_stream_bound"] = sub_modes_banded(sub_steps)
        torch.cuda.empty_cache()
        work["spmv_powerlaw_4m_f64"] = sub_spmv_f64(sub_steps)
        torch.cuda.empty_cache()
        work["spmm_banded_1m"] = sub_spmm_c3(sub_steps, args.no_vendor)
        torch.cuda.empty_cache()
        work["batched_512"] = sub_batched(sub_steps)
        work["seconds"] = round(time.time() - t_sub, 1)
        result["workloads"] = work
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
