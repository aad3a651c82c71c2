"""CSR SpMV on gfx950 behind the reference's operator interface.

Same public names, keyword defaults, exception types and messages as the
reference's ``flagsparse/sparse_operations/spmv_csr.py`` (cited per function
below); the kernels are the hand-written HIP ones of ``csrc/spmv.hip``,
reached through the C ABI (``_lib``).

Differences that keep results identical:
  * the default path widens fp32 to fp64 inside the kernel instead of casting
    ``data`` and ``x`` to fp64 tensors first (spmv_csr.py:652-689);
  * validation runs as one device pass (``fs_csr_analyze``) with one host copy
    instead of one host synchronisation per check (spmv_csr.py:525-542);
  * ``use_opt=True`` runs the same row-split kernel with native accumulation
    instead of one launch per row-length bucket (spmv_csr.py:393-440).
Extensions: trailing ``alpha`` / ``beta`` keywords (defaults 1 / 0 give the
reference's write-only behaviour) and an optional ``schedule``.
"""

from __future__ import annotations

import time

import torch

from .. import _lib
from ..rowsplit import RowSplitPlan, Schedule, default_spmv_schedule, supported_spmv_schedule
from ..schedule_choice import choose_schedule, exhaustive_spmv_schedule
from ..slicing import SLICED_DTYPES, SlicedCsr
from ..sortkey import stable_order
from ..xwin import XWIN_DTYPES, XWinCsr
from ._common import (
    SUPPORTED_INDEX_DTYPES,
    _INDEX_LIMIT_INT32,
    _is_complex_dtype,
)

SUPPORTED_SPMV_VALUE_DTYPES = (
    torch.float16,
    torch.bfloat16,
    torch.float32,
    torch.float64,
    torch.complex64,
    torch.complex128,
)

SPMV_OP_NON = 0
SPMV_OP_TRANS = 1
SPMV_OP_CONJ_TRANS = 2
SPMV_OP_NAMES = {SPMV_OP_NON: "non", SPMV_OP_TRANS: "trans", SPMV_OP_CONJ_TRANS: "conj"}
_SPMV_OP_NAME_TO_CODE = {v: k for k, v in SPMV_OP_NAMES.items()}
_OP_ERROR = "op must be one of: 0=non, 1=trans, 2=conj"


def _normalize_spmv_op(op=None, transpose=False):
    """op code from int / name / transpose flag (spmv_csr.py:29-44)."""
    if op is None:
        return SPMV_OP_TRANS if bool(transpose) else SPMV_OP_NON
    if isinstance(op, str):
        code = _SPMV_OP_NAME_TO_CODE.get(op.strip().lower())
        if code is None:
            raise ValueError(_OP_ERROR)
        return code
    try:
        code = int(op)
    except (TypeError, ValueError) as exc:
        raise ValueError(_OP_ERROR) from exc
    if code not in SPMV_OP_NAMES:
        raise ValueError(_OP_ERROR)
    return code


def _spmv_op_to_name(op):
    return SPMV_OP_NAMES[_normalize_spmv_op(op)]


def _spmv_op_transposes(op):
    return _normalize_spmv_op(op) in (SPMV_OP_TRANS, SPMV_OP_CONJ_TRANS)


class PreparedCsrSpmv:
    """Cached CSR metadata for repeated SpMV calls (spmv_csr.py:55-122).

    Holds the same fields as the reference.  ``opt_buckets`` holds the
    reference's row-length bucket lists (``_build_spmv_opt_buckets``); when
    the constructor gets ``opt_buckets=None`` they are built on first read,
    so calls that never read them pay no host synchronisation for them.
    ``schedule`` is the HIP schedule that actually runs and
    ``schedule_record`` says how it was chosen.  Device plans are built on
    first use and cached per accumulation mode, so a directly constructed
    instance works too.
    """

    __slots__ = (
        "data", "kernel_indices", "kernel_indptr", "shape", "n_rows", "n_cols",
        "block_nnz", "max_segments", "max_row_nnz", "_opt_buckets", "supports_opt",
        "transpose", "op", "index_fallback_policy", "index_fallback_applied",
        "index_fallback_reason", "_baseline_compute_dtype", "_baseline_data",
        "schedule", "_plans", "_csr", "row_lengths", "schedule_record",
        "_bucket_max_segments", "_stats",
    )

    def __init__(self, data, kernel_indices, kernel_indptr, shape, n_rows, n_cols, block_nnz,
                 max_segments, max_row_nnz, opt_buckets, transpose=False, op=None,
                 index_fallback_policy="auto", index_fallback_applied=False,
                 index_fallback_reason=None, schedule=None, row_lengths=None):
        self.data = data
        self.kernel_indices = kernel_indices
        self.kernel_indptr = kernel_indptr
        self.shape = (int(shape[0]), int(shape[1]))
        self.n_rows = n_rows
        self.n_cols = n_cols
        self.block_nnz = block_nnz
        self.max_segments = max_segments
        self.max_row_nnz = max_row_nnz
        self._opt_buckets = opt_buckets
        self._bucket_max_segments = None
        self.schedule_record = None
        self.supports_opt = (data.dtype in (torch.float32, torch.float64)
                             and kernel_indices.dtype == torch.int32)
        self.op = _normalize_spmv_op(op, transpose=transpose)
        self.transpose = _spmv_op_transposes(self.op)
        self.index_fallback_policy = str(index_fallback_policy).lower()
        self.index_fallback_applied = bool(index_fallback_applied)
        self.index_fallback_reason = index_fallback_reason
        if data.dtype in (torch.float16, torch.bfloat16):
            self._baseline_compute_dtype = torch.float32
        elif data.dtype == torch.float32:
            self._baseline_compute_dtype = torch.float64
        else:
            self._baseline_compute_dtype = data.dtype
        self._baseline_data = None
        self.schedule = schedule
        self.row_lengths = row_lengths
        self._plans = {}
        self._csr = None
        self._stats = None  # validation statistics of the caller's matrix (prepare_spmv_csr)

    @property
    def opt_buckets(self):
        """Row lists per length bucket, as spmv_csr.py:345-391 builds them."""
        if self._opt_buckets is None:
            lens = self.row_lengths
            if lens is None:
                lens = (self.kernel_indptr[1:] - self.kernel_indptr[:-1]).to(torch.int64)
            row_index_dtype = torch.int32 if self.n_rows <= _INDEX_LIMIT_INT32 else torch.int64
            self._opt_buckets = _build_spmv_opt_buckets(
                lens, max_row_nnz=self.max_row_nnz, row_index_dtype=row_index_dtype,
                max_segments=self._bucket_max_segments,
                fp64=self.data.dtype == torch.float64)
        return self._opt_buckets

    @opt_buckets.setter
    def opt_buckets(self, value):
        self._opt_buckets = value

    # -- device plumbing (not part of the reference surface) --------------
    def csr_struct(self):
        if self._csr is None:
            self._csr = _lib.make_csr(self.data, self.kernel_indices, self.kernel_indptr,
                                      self.n_rows, self.n_cols)
        return self._csr

    def plan(self, accumulate, schedule=None):
        sch = schedule or self.schedule or default_spmv_schedule(self.n_rows, self.data.numel(),
                                                                 self.data.dtype, self.n_cols)
        key = (int(accumulate), sch.key())
        p = self._plans.get(key)
        if p is None:
            p = RowSplitPlan(self.kernel_indptr, self.n_rows, self.data.dtype, sch, accumulate,
                             row_lengths=self.row_lengths)
            self._plans[key] = p
        return p


# Reference row-length bucket tables (spmv_csr.py:128-153), kept for
# _build_spmv_opt_buckets; the HIP path schedules by Schedule instead.
_SPMV_OPT_BUCKET_CONFIGS = (
    {"max_row_nnz": 64, "block_size": 32, "num_warps": 1, "num_stages": 2, "batch_rows": 16},
    {"max_row_nnz": 512, "block_size": 256, "num_warps": 8, "num_stages": 2},
    {"max_row_nnz": 4096, "block_size": 512, "num_warps": 16, "num_stages": 2},
    {"max_row_nnz": None, "block_size": 1024, "num_warps": 32, "num_stages": 3},
)
_SPMV_OPT_BUCKET_CONFIGS_FP64 = (
    {"max_row_nnz": 64, "block_size": 32, "num_warps": 1, "num_stages": 2, "batch_rows": 4},
    {"max_row_nnz": 256, "block_size": 64, "num_warps": 2, "num_stages": 2},
    {"max_row_nnz": 2048, "block_size": 128, "num_warps": 4, "num_stages": 2},
    {"max_row_nnz": 8192, "block_size": 256, "num_warps": 8, "num_stages": 2},
    {"max_row_nnz": None, "block_size": 512, "num_warps": 16, "num_stages": 1},
)
_SPMV_OPT_ACC_MODES = ("fast", "mixed", "accurate")


def _build_spmv_opt_buckets(row_lengths, max_row_nnz, row_index_dtype, max_segments=None,
                            fp64=False):
    """Row lists per length bucket (spmv_csr.py:345-391): first bucket takes
    lengths <= its bound including empty rows, the last takes the rest."""
    buckets = []
    lower = 0
    for spec in (_SPMV_OPT_BUCKET_CONFIGS_FP64 if fp64 else _SPMV_OPT_BUCKET_CONFIGS):
        upper = spec["max_row_nnz"]
        if upper is None:
            mask, bmax = row_lengths > lower, max_row_nnz
        elif lower == 0:
            mask, bmax = row_lengths <= upper, upper
        else:
            mask, bmax = (row_lengths > lower) & (row_lengths <= upper), upper
        rows = torch.nonzero(mask, as_tuple=False).flatten()
        if rows.numel() > 0:
            segs = max_segments if max_segments is not None else max(
                (bmax + spec["block_size"] - 1) // spec["block_size"], 1)
            buckets.append({
                "rows": rows.to(row_index_dtype),
                "block_size": spec["block_size"],
                "max_segs": segs,
                "num_warps": spec["num_warps"],
                "num_stages": spec["num_stages"],
                "batch_rows": int(spec.get("batch_rows", 1)),
            })
        if upper is not None:
            lower = upper
    return buckets


def _schedule_record(sch, how):
    """How the HIP schedule of a prepared matrix was chosen (not a reference
    field; the reference's ``opt_buckets`` keep their meaning)."""
    return {"schedule": sch, "chosen_by": how}


def _normalize_spmv_index_fallback_policy(index_fallback_policy):
    policy = str(index_fallback_policy).lower()
    if policy not in ("auto", "strict"):
        raise ValueError("index_fallback_policy must be 'auto' or 'strict'")
    return policy


def _spmv_dtype_error_message():
    return "data dtype must be one of: " + ", ".join(
        str(d).replace("torch.", "") for d in SUPPORTED_SPMV_VALUE_DTYPES)


def _transpose_csr_for_spmv(data, indices, indptr, shape):
    """CSR of A^T (spmv_csr.py:456-494): entries ordered by column with a
    stable sort, so equal columns keep row order; index dtype kept unless the
    new column count (old rows) or nnz exceeds int32."""
    n_rows, n_cols = int(shape[0]), int(shape[1])
    nnz = data.numel()
    device = data.device
    out_index_dtype = indices.dtype if n_rows <= _INDEX_LIMIT_INT32 else torch.int64
    out_indptr_dtype = indptr.dtype if nnz <= _INDEX_LIMIT_INT32 else torch.int64
    if nnz == 0:
        return (data, torch.empty(0, dtype=out_index_dtype, device=device),
                torch.zeros(n_cols + 1, dtype=out_indptr_dtype, device=device), (n_cols, n_rows))
    counts = (indptr[1:] - indptr[:-1]).to(torch.int64)
    row_of = torch.repeat_interleave(torch.arange(n_rows, dtype=torch.int64, device=device), counts)
    cols = indices.to(torch.int64)
    order = stable_order(cols, n_cols)
    t_indptr = torch.zeros(n_cols + 1, dtype=torch.int64, device=device)
    t_indptr[1:] = torch.cumsum(torch.bincount(cols, minlength=n_cols), dim=0)
    return (data[order].contiguous(), row_of[order].to(out_index_dtype).contiguous(),
            t_indptr.to(out_indptr_dtype).contiguous(), (n_cols, n_rows))


def _prepare_spmv_csr_matrix(data, indices, indptr, shape, index_fallback_policy="auto",
                             _stats_out=None, _lengths=True):
    """Validation with the reference's checks, order and messages
    (spmv_csr.py:497-551); the device-side checks run as one fs_csr_analyze
    pass.  Returns (data, indices, indptr, n_rows, n_cols, row_lengths,
    max_row_nnz); row_lengths is None when ``_lengths`` is false (a
    validation-only pass)."""
    _normalize_spmv_index_fallback_policy(index_fallback_policy)
    if not all(torch.is_tensor(t) for t in (data, indices, indptr)):
        raise TypeError("data, indices, indptr must all be torch.Tensor")
    if data.ndim != 1 or indices.ndim != 1 or indptr.ndim != 1:
        raise ValueError("data, indices, indptr must be 1D tensors")
    n_rows, n_cols = int(shape[0]), int(shape[1])
    if indptr.numel() != n_rows + 1:
        raise ValueError(f"indptr length must be n_rows+1={n_rows + 1}, got {indptr.numel()}")
    if data.numel() != indices.numel():
        raise ValueError("data and indices must have the same length (nnz)")
    if not all(t.is_cuda for t in (data, indices, indptr)):
        raise ValueError("data, indices, indptr must be CUDA tensors")
    if not all(t.device == data.device for t in (indices, indptr)):
        raise ValueError("data, indices, indptr must be on the same CUDA device")
    if data.dtype not in SUPPORTED_SPMV_VALUE_DTYPES:
        raise TypeError(_spmv_dtype_error_message())
    if indices.dtype not in SUPPORTED_INDEX_DTYPES:
        raise TypeError("indices dtype must be torch.int32 or torch.int64")
    if indptr.dtype not in SUPPORTED_INDEX_DTYPES:
        raise TypeError("indptr dtype must be torch.int32 or torch.int64")
    data = data.contiguous()
    indices = indices.contiguous()
    indptr = indptr.contiguous()
    nnz = data.numel()
    stats_dev = _lib.analyze_device(_lib.make_csr(data, indices, indptr, n_rows, n_cols), data.device)
    stats = stats_dev.cpu().tolist()
    if _stats_out is not None:
        _stats_out.append((stats_dev, stats))
    if stats[0] != 0:
        raise ValueError("indptr must start at zero")
    if stats[1] != nnz:
        raise ValueError("indptr[-1] must equal nnz")
    if stats[2] > 0:
        raise ValueError("indptr must be non-decreasing")
    if nnz > 0 and (stats[3] < 0 or stats[4] >= n_cols):
        raise IndexError("indices out of range for n_cols")
    row_lengths = (indptr[1:] - indptr[:-1]).to(torch.int64) if _lengths else None
    max_row_nnz = int(stats[5]) if n_rows > 0 else 0
    return data, indices, indptr, n_rows, n_cols, row_lengths, max_row_nnz


def _validate_spmv_x(x, prepared):
    """x checks of spmv_csr.py:554-567."""
    if x is None or not torch.is_tensor(x):
        raise TypeError("x must be a torch.Tensor")
    if x.ndim != 1:
        raise ValueError("x must be a 1D tensor")
    if not x.is_cuda:
        raise ValueError("x must be a CUDA tensor")
    if x.dtype != prepared.data.dtype:
        raise TypeError("x dtype must match sparse matrix dtype")
    if x.numel() != prepared.n_cols:
        raise ValueError(f"x length must be n_cols={prepared.n_cols}, got {x.numel()}")
    if x.device != prepared.data.device:
        raise ValueError("x must be on the same device as sparse matrix data")
    return x.contiguous()


def prepare_spmv_csr(data, indices, indptr, shape, block_nnz=256, max_segments=None,
                     transpose=False, op=None, index_fallback_policy="auto", schedule=None):
    """Validate once and cache metadata (spmv_csr.py:570-649).

    ``block_nnz`` / ``max_segments`` are validated and recorded with the
    reference's auto-doubling rule (:613-626) for API compatibility.  The HIP
    schedule is ``schedule``: a ``Schedule``, "learned" (the packaged
    selector, run on the device; the default when a model is installed),
    "exhaustive" (time every variant now) or "default" (the fixed rule).
    """
    index_fallback_policy = _normalize_spmv_index_fallback_policy(index_fallback_policy)
    op_code = _normalize_spmv_op(op, transpose=transpose)
    if op is not None and bool(transpose) and op_code == SPMV_OP_NON:
        raise ValueError("transpose=True conflicts with op=non")
    transpose = _spmv_op_transposes(op_code)
    caller_stats = []
    if transpose:
        data, indices, indptr, *_ = _prepare_spmv_csr_matrix(
            data, indices, indptr, shape, index_fallback_policy=index_fallback_policy,
            _stats_out=caller_stats)
        if op_code == SPMV_OP_CONJ_TRANS and _is_complex_dtype(data.dtype):
            data = data.conj().resolve_conj()
        data, indices, indptr, shape = _transpose_csr_for_spmv(data, indices, indptr, shape)
    stats_out = []
    (data, kernel_indices, kernel_indptr, n_rows, n_cols, row_lengths,
     max_row_nnz) = _prepare_spmv_csr_matrix(data, indices, indptr, shape,
                                             index_fallback_policy=index_fallback_policy,
                                             _stats_out=stats_out)
    block_nnz_use = block_nnz
    if max_segments is None:
        max_segments_use = max((max_row_nnz + block_nnz_use - 1) // block_nnz_use, 1)
        while max_segments_use > 2048 and block_nnz_use < 65536:
            block_nnz_use *= 2
            max_segments_use = max((max_row_nnz + block_nnz_use - 1) // block_nnz_use, 1)
    else:
        max_segments_use = max_segments
    sch, how = choose_schedule("spmv", schedule, data, n_rows, n_cols, stats_out[0][0],
                               kernel_indptr, kernel_indices)
    if how != "given":  # a library choice runs as what is built for these dtypes
        sch = supported_spmv_schedule(sch, data.dtype, kernel_indices.dtype, data.numel())
    prepared = PreparedCsrSpmv(
        data=data, kernel_indices=kernel_indices, kernel_indptr=kernel_indptr, shape=shape,
        n_rows=n_rows, n_cols=n_cols, block_nnz=block_nnz_use, max_segments=max_segments_use,
        max_row_nnz=max_row_nnz, opt_buckets=None, transpose=transpose, op=op_code,
        index_fallback_policy=index_fallback_policy, schedule=sch, row_lengths=row_lengths)
    prepared._bucket_max_segments = max_segments
    if how == "exhaustive":
        sch = exhaustive_spmv_schedule(prepared)
        prepared.schedule = sch
    prepared.schedule_record = _schedule_record(sch, how)
    prepared._stats = (caller_stats or stats_out)[0][1]
    return prepared


class _UnpreparedCache:
    """Prepared objects of recent unprepared calls (``flagsparse_spmv_csr``
    given data / indices / indptr and no ``prepared``), so a repeated call
    on the same matrix does not rebuild its schedule, layouts and selection.

    An entry is reused only for the very same tensor objects (held by weak
    reference) with unchanged version counters (every in-place write through
    torch bumps them) and data pointers, the same shape and options.  Every
    reuse still runs the validation pass on the caller's tensors (one
    fs_csr_analyze and its host read) and raises exactly as preparing again
    would; if that pass's structure statistics (row-length histogram, column
    range, ...) differ from the cached ones, the matrix is prepared again.
    """

    def __init__(self, size=4):
        self.size = size
        self._entries = []  # (key, weakrefs, prepared), most recent last

    @staticmethod
    def _key(tensors, shape, options):
        return (tuple((t.data_ptr(), t._version, t.dtype, tuple(t.shape), t.stride())
                      for t in tensors), (int(shape[0]), int(shape[1])), options)

    def lookup(self, tensors, shape, options):
        key = self._key(tensors, shape, options)
        for i, (k, refs, prepared) in enumerate(self._entries):
            if k == key and all(r() is t for r, t in zip(refs, tensors)):
                self._entries.append(self._entries.pop(i))
                return prepared
        return None

    def store(self, tensors, shape, options, prepared):
        import weakref
        key = self._key(tensors, shape, options)
        # drop entries of the same key and stale ones for the same tensor
        # objects (their contents have changed since)
        self._entries = [e for e in self._entries
                         if e[0] != key and not (e[0][1:] == key[1:]
                                                 and all(r() is t for r, t in zip(e[1], tensors)))]
        self._entries.append((key, [weakref.ref(t) for t in tensors], prepared))
        del self._entries[:-self.size]

    def clear(self):
        self._entries = []


_UNPREPARED = _UnpreparedCache()


def _prepare_for_call(data, indices, indptr, shape, block_nnz, max_segments, op_code,
                      index_fallback_policy):
    """The prepared object of an unprepared call: a cached one after a fresh
    validation pass (see _UnpreparedCache), else prepare_spmv_csr."""
    tensors = (data, indices, indptr)
    options = (block_nnz, max_segments, op_code, index_fallback_policy)
    if all(torch.is_tensor(t) for t in tensors):
        prepared = _UNPREPARED.lookup(tensors, shape, options)
        if prepared is not None:
            stats = []
            _prepare_spmv_csr_matrix(data, indices, indptr, shape,
                                     index_fallback_policy=index_fallback_policy,
                                     _stats_out=stats, _lengths=False)
            if stats[0][1] == prepared._stats:
                return prepared
    prepared = prepare_spmv_csr(data, indices, indptr, shape, block_nnz=block_nnz,
                                max_segments=max_segments, op=op_code,
                                index_fallback_policy=index_fallback_policy)
    _UNPREPARED.store(tensors, shape, options, prepared)
    return prepared


def _get_spmv_baseline_data(prepared):
    """(compute dtype, data) of the default path.  The HIP kernel widens in
    registers, so no widened copy of ``data`` is made (cf. spmv_csr.py:652-661)."""
    return prepared._baseline_compute_dtype, prepared.data


# fp32 values on the default path sum fp64 products: 8 bytes of LDS per
# entry, so a 4096-entry tile leaves 4 blocks per CU.  Their x-window layouts
# use tiles of at most this many entries (8 blocks per CU): on the 4M-row
# power-law matrix the default path takes 181 instead of 205 us per call,
# with the same result bits.
XWIN_WIDE_TILE = 2048


def xwin_plan_key(prepared, sch, accumulate):
    """Key of the x-window layout in ``prepared._plans`` for schedule ``sch``
    and accumulation mode ``accumulate``, and the layout's tile size."""
    T = sch.tile_nnz if sch.tile_nnz > 0 else 2048
    if accumulate == _lib.FS_ACC_WIDE and prepared.data.dtype == torch.float32:
        T = min(T, XWIN_WIDE_TILE)
    return ("xwin", sch.key(), T), T


def _launch_spmv(prepared, x, accumulate, out=None, alpha=1.0, beta=0.0, schedule=None):
    y = out if out is not None else torch.empty(prepared.n_rows, dtype=prepared.data.dtype,
                                                 device=prepared.data.device)
    if prepared.n_rows == 0:
        return y
    sch = schedule or prepared.schedule
    if (sch is not None and sch.x_window != 0 and prepared.data.dtype in XWIN_DTYPES
            and 0 < prepared.data.numel() <= 2**31 - 1):
        # two streaming passes: x gathered from LDS windows, then entry tiles
        key, T = xwin_plan_key(prepared, sch, accumulate)
        xw = prepared._plans.get(key)
        if xw is None:
            xw = XWinCsr(prepared.data, prepared.kernel_indices, prepared.kernel_indptr,
                         prepared.n_rows, prepared.n_cols, tile_nnz=T,
                         window=sch.x_window if sch.x_window > 0 else None)
            prepared._plans[key] = xw
        xw.run(x, y, alpha, beta, accumulate, prepared.data.device)
        return y
    if (sch is not None and sch.x_slice_cols > 0 and prepared.n_cols > sch.x_slice_cols
            and prepared.data.dtype in SLICED_DTYPES):
        # native partials in the value dtype; the default fp32 path (fp64
        # accumulation, one rounding) keeps raw fp64 row sums as partials
        key = ("sliced", sch.key(), int(accumulate))
        sliced = prepared._plans.get(key)
        if sliced is None:
            sliced = SlicedCsr(prepared.data, prepared.kernel_indices, prepared.kernel_indptr,
                               prepared.n_rows, prepared.n_cols, sch.x_slice_cols, sch,
                               accumulate, sums=accumulate == _lib.FS_ACC_WIDE)
            prepared._plans[key] = sliced
        sliced.run(x, y, alpha, beta, prepared.data.device)
        return y
    plan = prepared.plan(accumulate, schedule)
    _lib.spmv(prepared.csr_struct(), x, y, alpha, beta, plan.struct(1), prepared.data.device)
    return y


def _hip_spmv_csr_impl_prepared(prepared, x, out=None, alpha=1.0, beta=0.0, schedule=None):
    """Default path: fp32 accumulated in fp64 and rounded once, fp16 / bf16 in
    fp32, fp64 / complex natively -- the numerics of spmv_csr.py:664-708."""
    acc = _lib.FS_ACC_WIDE if prepared.data.dtype == torch.float32 else _lib.FS_ACC_NATIVE
    return _launch_spmv(prepared, x, acc, out, alpha, beta, schedule)


def _hip_spmv_csr_impl_opt_prepared(prepared, x, out=None, alpha=1.0, beta=0.0, schedule=None):
    """use_opt path: native fp32 / fp64 accumulation (spmv_csr.py:393-440);
    chunked long rows of fp32 use compensated summation."""
    return _launch_spmv(prepared, x, _lib.FS_ACC_NATIVE, out, alpha, beta, schedule)


# The reference's launcher names, kept so code (and tests) that patch or call
# them by name keep working.
_triton_spmv_csr_impl_prepared = _hip_spmv_csr_impl_prepared
_triton_spmv_csr_impl_opt_prepared = _hip_spmv_csr_impl_opt_prepared


def _spmv_uses_int64_indices(prepared):
    return (prepared.kernel_indices.dtype == torch.int64
            or prepared.kernel_indptr.dtype == torch.int64)


def _spmv_int32_fallback_blocker(prepared):
    """Why an int32 retry is unsafe, or None (spmv_csr.py:718-735)."""
    if prepared.kernel_indices.dtype == torch.int64 and prepared.kernel_indices.numel() > 0:
        lo = int(prepared.kernel_indices.min().item())
        hi = int(prepared.kernel_indices.max().item())
        if lo < 0 or hi > _INDEX_LIMIT_INT32:
            return (f"column index range [{lo}, {hi}] cannot fit int32 "
                    f"for shape={prepared.shape}")
    if prepared.kernel_indptr.dtype == torch.int64 and prepared.kernel_indptr.numel() > 0:
        last = int(prepared.kernel_indptr[-1].item())
        if last > _INDEX_LIMIT_INT32:
            return f"CSR nnz offset {last} cannot fit int32 for shape={prepared.shape}"
    if prepared.n_rows > _INDEX_LIMIT_INT32:
        return f"row count {prepared.n_rows} cannot fit int32 row metadata"
    return None


def _spmv_prepared_with_int32_indices(prepared, reason):
    """int32 copy of a prepared int64 matrix for the retry (spmv_csr.py:738-774)."""
    blocker = _spmv_int32_fallback_blocker(prepared)
    if blocker is not None:
        raise RuntimeError(f"native int64 CSR SpMV failed and int32 fallback is unsafe: {blocker}")
    kernel_indices = prepared.kernel_indices.to(torch.int32)
    kernel_indptr = prepared.kernel_indptr.to(torch.int32)
    row_lengths = (kernel_indptr[1:] - kernel_indptr[:-1]).to(torch.int64)
    max_row_nnz = int(row_lengths.max().item()) if prepared.n_rows > 0 else 0
    fb = PreparedCsrSpmv(
        data=prepared.data, kernel_indices=kernel_indices, kernel_indptr=kernel_indptr,
        shape=prepared.shape, n_rows=prepared.n_rows, n_cols=prepared.n_cols,
        block_nnz=prepared.block_nnz, max_segments=prepared.max_segments,
        max_row_nnz=max_row_nnz, opt_buckets=None, transpose=prepared.transpose, op=prepared.op,
        index_fallback_policy=prepared.index_fallback_policy, index_fallback_applied=True,
        index_fallback_reason=str(reason), schedule=prepared.schedule, row_lengths=row_lengths)
    fb._bucket_max_segments = prepared.max_segments  # as spmv_csr.py:751-757 rebuilds them
    fb.schedule_record = prepared.schedule_record
    return fb


def _launch_kwargs(out, alpha, beta):
    """Keyword arguments for a launcher hook: none with the reference's
    defaults, so hooks patched with the reference's two-argument signature
    keep working."""
    if out is None and float(alpha) == 1.0 and float(beta) == 0.0:
        return {}
    return {"out": out, "alpha": alpha, "beta": beta}


def _run_spmv_prepared(prepared, x, use_opt=False, out=None, alpha=1.0, beta=0.0):
    """One launch (spmv_csr.py:777-780); ``out`` / ``alpha`` / ``beta`` are
    the trailing extension (y = alpha A x + beta out, written into ``out``)."""
    kw = _launch_kwargs(out, alpha, beta)
    if use_opt and prepared.supports_opt:
        return _triton_spmv_csr_impl_opt_prepared(prepared, x, **kw)
    return _triton_spmv_csr_impl_prepared(prepared, x, **kw)


def _run_spmv_prepared_with_fallback(prepared, x, use_opt=False, out=None, alpha=1.0, beta=0.0):
    """Retry with int32 indices after a failed int64 launch under policy
    'auto' (spmv_csr.py:783-793).  Launch failures are raised before any
    kernel runs, so a retry with beta != 0 still reads the caller's y."""
    kw = _launch_kwargs(out, alpha, beta)
    try:
        return _run_spmv_prepared(prepared, x, use_opt=use_opt, **kw)
    except Exception as exc:
        if prepared.index_fallback_policy != "auto" or not _spmv_uses_int64_indices(prepared):
            raise
        fb = _spmv_prepared_with_int32_indices(prepared, exc)
        return _run_spmv_prepared(fb, x, use_opt=use_opt, **kw)


def flagsparse_spmv_csr(data=None, indices=None, indptr=None, x=None, shape=None,
                        block_nnz=256, max_segments=None, out=None, return_time=False,
                        use_opt=False, prepared=None, transpose=None, op=None,
                        index_fallback_policy="auto", alpha=1.0, beta=0.0):
    """y = op(A) x (spmv_csr.py:796-872); op 0/'non', 1/'trans', 2/'conj'.

    ``alpha`` / ``beta`` extend the reference: y = alpha op(A) x + beta y,
    where y is ``out`` (required when beta != 0).
    """
    op_explicit = op is not None
    op_code = _normalize_spmv_op(op, transpose=False if transpose is None else bool(transpose))
    if op_explicit and transpose is not None and bool(transpose) != _spmv_op_transposes(op_code):
        raise ValueError("transpose conflicts with op")
    if prepared is None:
        if any(a is None for a in (data, indices, indptr, shape)):
            raise ValueError(
                "data, indices, indptr, and shape are required when prepared is not provided")
        prepared = _prepare_for_call(data, indices, indptr, shape, block_nnz, max_segments,
                                     op_code, index_fallback_policy)
    else:
        if op_explicit and op_code != prepared.op:
            raise ValueError(f"op={_spmv_op_to_name(op_code)} does not match "
                             f"prepared.op={_spmv_op_to_name(prepared.op)}")
        if not op_explicit and transpose is not None and bool(transpose) != prepared.transpose:
            raise ValueError(f"transpose={bool(transpose)} does not match "
                             f"prepared.transpose={prepared.transpose}")
    x = _validate_spmv_x(x, prepared)
    if out is not None:
        if not out.is_cuda:
            raise ValueError("out must be a CUDA tensor")
        if out.device != prepared.data.device:
            raise ValueError("out must be on the same CUDA device as the result")
        if tuple(out.shape) != (prepared.n_rows,) or out.dtype != prepared.data.dtype:
            raise ValueError("out shape/dtype must match result")
    if float(beta) != 0.0 and out is None:
        raise ValueError("beta != 0 requires out (the y that is scaled and accumulated)")
    # with a contiguous `out` the kernels write it directly (the reference
    # computes a new y and copies it into `out`: same values, one pass fewer);
    # a strided `out` is staged through a contiguous copy when beta reads it
    if out is None or out.is_contiguous():
        target = out
    else:
        target = out.contiguous() if float(beta) != 0.0 else None
    t0 = None
    if return_time:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
    y = _run_spmv_prepared_with_fallback(prepared, x, use_opt=use_opt, out=target, alpha=alpha,
                                         beta=beta)
    elapsed_ms = None
    if return_time:
        torch.cuda.synchronize()
        elapsed_ms = (time.perf_counter() - t0) * 1000.0
    if out is not None and y is not out:
        out.copy_(y)
        y = out
    if return_time:
        return y, elapsed_ms
    return y


# ---------------------------------------------------------------- COO -> CSR

def _coo_is_sorted_lex(row_i64, col_i64, n_cols):
    """True iff (row, col) pairs are non-decreasing lexicographically."""
    n = row_i64.numel()
    if n <= 1:
        return True
    key = row_i64 * max(1, int(n_cols)) + col_i64
    return bool((key[1:] >= key[:-1]).all().item())


def coo_to_csr_for_spmv(data, row, col, shape, assume_sorted=False):
    """COO -> (data, int32 columns, int64 indptr) sorted by (row, col)
    (spmv_csr.py:885-912).  The sort is stable (the reference's argsort is
    not), so repeated (row, col) pairs keep their input order and are summed
    in that order: the result is reproducible bit for bit."""
    n_rows, n_cols = int(shape[0]), int(shape[1])
    row64, col64 = row.to(torch.int64), col.to(torch.int64)
    if row64.numel() == 0:
        return data, col64.to(torch.int32), torch.zeros(n_rows + 1, dtype=torch.int64,
                                                        device=data.device)
    if assume_sorted or _coo_is_sorted_lex(row64, col64, n_cols):
        row_s, col_s, data_s = row64, col64, data
    else:
        order = torch.argsort(row64 * max(1, n_cols) + col64, stable=True)
        row_s, col_s, data_s = row64[order], col64[order], data[order]
    indptr = torch.zeros(n_rows + 1, dtype=torch.int64, device=data.device)
    indptr[1:] = torch.cumsum(torch.bincount(row_s, minlength=n_rows), dim=0)
    return data_s, col_s.to(torch.int32), indptr


def prepare_spmv_coo_tocsr(data, row, col, shape, block_nnz=256, max_segments=None,
                           assume_sorted=False):
    """One-time COO -> CSR + SpMV metadata (spmv_csr.py:915-946)."""
    if not all(torch.is_tensor(t) for t in (data, row, col)):
        raise TypeError("data, row, col must all be torch.Tensor")
    if not all(t.is_cuda for t in (data, row, col)):
        raise ValueError("data, row, col must all be CUDA tensors")
    if data.ndim != 1 or row.ndim != 1 or col.ndim != 1:
        raise ValueError("data, row, col must all be 1D tensors")
    if data.dtype not in SUPPORTED_SPMV_VALUE_DTYPES:
        raise TypeError(_spmv_dtype_error_message())
    if row.numel() != col.numel() or data.numel() != row.numel():
        raise ValueError("data, row, col must have the same length")
    data_s, indices, indptr = coo_to_csr_for_spmv(data, row, col, shape,
                                                  assume_sorted=assume_sorted)
    return prepare_spmv_csr(data_s, indices, indptr, shape, block_nnz=block_nnz,
                            max_segments=max_segments)


def flagsparse_spmv_coo_tocsr(data=None, row=None, col=None, x=None, shape=None, block_nnz=256,
                              max_segments=None, out=None, return_time=False, use_opt=True,
                              prepared=None, assume_sorted=False):
    """COO SpMV through the CSR path (spmv_csr.py:949-1016)."""
    if prepared is not None:
        if x is None:
            raise TypeError("x is required")
        if shape is None:
            shape = prepared.shape
        sh = (int(shape[0]), int(shape[1]))
        if sh != prepared.shape:
            raise ValueError(f"shape {sh} does not match prepared.shape {prepared.shape}")
        return flagsparse_spmv_csr(x=x, shape=shape, block_nnz=block_nnz,
                                   max_segments=max_segments, out=out, return_time=return_time,
                                   use_opt=use_opt, prepared=prepared)
    if not all(torch.is_tensor(t) for t in (data, row, col, x)):
        raise TypeError("data, row, col, x must all be torch.Tensor")
    if not all(t.is_cuda for t in (data, row, col, x)):
        raise ValueError("data, row, col, x must all be CUDA tensors")
    if data.ndim != 1 or row.ndim != 1 or col.ndim != 1 or x.ndim != 1:
        raise ValueError("data, row, col, x must all be 1D tensors")
    if data.dtype not in SUPPORTED_SPMV_VALUE_DTYPES:
        raise TypeError(_spmv_dtype_error_message())
    if x.dtype != data.dtype:
        raise TypeError("x dtype must match data dtype")
    data_s, indices, indptr = coo_to_csr_for_spmv(data, row, col, shape,
                                                  assume_sorted=assume_sorted)
    return flagsparse_spmv_csr(data_s, indices, indptr, x, shape, block_nnz=block_nnz,
                               max_segments=max_segments, out=out, return_time=return_time,
                               use_opt=use_opt)


__all__ = [
    "SUPPORTED_SPMV_VALUE_DTYPES",
    "SPMV_OP_NON",
    "SPMV_OP_TRANS",
    "SPMV_OP_CONJ_TRANS",
    "PreparedCsrSpmv",
    "prepare_spmv_csr",
    "flagsparse_spmv_csr",
    "coo_to_csr_for_spmv",
    "prepare_spmv_coo_tocsr",
    "flagsparse_spmv_coo_tocsr",
    "Schedule",
]
