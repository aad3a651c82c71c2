"""How a prepared matrix gets its row-split schedule.

``choose_schedule(op, request, ...)`` resolves a request to a ``Schedule``:
  * a ``Schedule`` instance is used as given;
  * "default" is the fixed rule of ``rowsplit``;
  * "learned" (also what None means) runs the packaged selector for ``op`` on
    the device -- features from the analysis pass's statistics, MLP forward,
    argmin -- and falls back to the fixed rule when no model is installed, the
    values are not fp32 (SpMV: fp32 or fp64; fp64 SpMV uses the model trained
    on the same matrices timed with fp64 values), or the matrix is larger than
    any the model was trained on;
  * "exhaustive" is resolved after the prepared object exists
    (``exhaustive_spmv_schedule`` times every variant).
Returns (schedule or None, how).
"""

from __future__ import annotations

import torch

from .rowsplit import Schedule, default_spmm_schedule, default_spmv_schedule

_MODES = ("learned", "default", "exhaustive")


def choose_schedule(op, request, data, n_rows, n_cols, stats_dev, indptr=None, indices=None):
    """``indptr`` / ``indices``: the matrix's row pointers and columns, which
    the SpMV fixed rule samples for column locality when x is large."""
    nnz = data.numel()
    if isinstance(request, Schedule):
        return request, "given"
    mode = "learned" if request is None else str(request)
    if mode not in _MODES:
        raise ValueError(f"schedule must be a Schedule or one of {_MODES}")

    def fallback():
        if op == "spmv":
            return default_spmv_schedule(n_rows, nnz, data.dtype, n_cols, indptr, indices)
        return default_spmm_schedule(n_rows, nnz, 64)
    if mode == "default" or n_rows == 0:
        return fallback(), "default"
    if mode == "exhaustive":
        if op != "spmv":
            raise ValueError("exhaustive schedule search is offered for SpMV")
        return fallback(), "exhaustive"
    if data.dtype != torch.float32 and not (op == "spmv" and data.dtype == torch.float64):
        return fallback(), "default"
    from .selector import load_default
    sel = None
    if op == "spmv" and data.dtype == torch.float64:
        # the model trained on fp64 timings where one is packaged; else the
        # fp32-trained one (its features describe only the sparsity structure)
        sel = load_default(op, data.device, "f64")
    if sel is None:
        sel = load_default(op, data.device)
    if sel is None:
        return fallback(), "default"
    if not sel.covers(n_rows, n_cols, nnz):
        # outside the sizes the network was trained on: the fixed rule
        return fallback(), "default (outside the selector's training range)"
    sch, idx = sel.schedule_for(stats_dev, n_rows, n_cols, nnz)
    return sch, f"learned:{idx}"


def exhaustive_spmv_schedule(prepared, reps=5):
    """Time every SpMV variant on this matrix (random x) and return the fastest."""
    from .autotune import SPMV_VARIANTS, best_schedule, tune_spmv
    g = torch.Generator(device=prepared.data.device)
    g.manual_seed(0)
    x = (torch.rand(prepared.n_cols, generator=g, device=prepared.data.device) * 2 - 1).to(
        prepared.data.dtype)
    from .rowsplit import supported_spmv_schedule
    times = tune_spmv(prepared, x, reps=reps)
    return supported_spmv_schedule(best_schedule(times, SPMV_VARIANTS)[1], prepared.data.dtype,
                                   prepared.kernel_indices.dtype, prepared.data.numel())
