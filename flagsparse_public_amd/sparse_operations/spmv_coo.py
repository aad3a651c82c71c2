"""COO SpMV behind the reference's operator interface (spmv_coo.py).

Same names, arguments, validation order, messages and return conventions as
the reference's ``flagsparse_spmv_coo`` / ``prepare_spmv_coo`` /
``PreparedCoo`` (spmv_coo.py:37-700).  The reference keeps sorted COO plus
row-run bounds ``seg_starts`` and launches a row-run Triton kernel (or, with
``sort_by_row=False``, an atomic-add kernel over the entries).  Here the
prepared object keeps the same sorted arrays and ``seg_starts``, plus the
row pointer they imply, and runs the gfx950 CSR SpMV (``fs_spmv_csr``,
native accumulation like the reference's COO kernels).  Both ``sort_by_row``
settings sort, so the result is deterministic; with ``sort_by_row=False`` the
reference's atomics give run-to-run bit differences, ours do not.
Repeated (row, col) pairs are kept and summed in their sorted order.
"""

from __future__ import annotations

import time

import torch

from ._common import SUPPORTED_INDEX_DTYPES, _INDEX_LIMIT_INT32, _is_complex_dtype

SUPPORTED_SPMV_COO_VALUE_DTYPES = (torch.float32, torch.float64, torch.complex64,
                                   torch.complex128)

SPMV_COO_OP_NON = 0
SPMV_COO_OP_TRANS = 1
SPMV_COO_OP_CONJ_TRANS = 2
SPMV_COO_OP_NAMES = {SPMV_COO_OP_NON: "non", SPMV_COO_OP_TRANS: "trans",
                     SPMV_COO_OP_CONJ_TRANS: "conj"}
_SPMV_COO_OP_NAME_TO_CODE = {name: code for code, name in SPMV_COO_OP_NAMES.items()}


def _spmv_coo_dtype_error_message():
    return "COO SpMV supports float32, float64, complex64, and complex128"


def _normalize_spmv_coo_op(op=None, transpose=False):
    """op code 0/1/2 from int or name (spmv_coo.py:41-55)."""
    if op is None:
        return SPMV_COO_OP_TRANS if bool(transpose) else SPMV_COO_OP_NON
    if isinstance(op, str):
        token = op.strip().lower()
        if token not in _SPMV_COO_OP_NAME_TO_CODE:
            raise ValueError("op must be one of: 0=non, 1=trans, 2=conj")
        return _SPMV_COO_OP_NAME_TO_CODE[token]
    try:
        code = int(op)
    except (TypeError, ValueError) as exc:
        raise ValueError("op must be one of: 0=non, 1=trans, 2=conj") from exc
    if code not in SPMV_COO_OP_NAMES:
        raise ValueError("op must be one of: 0=non, 1=trans, 2=conj")
    return code


def _spmv_coo_op_to_name(op):
    return SPMV_COO_OP_NAMES[_normalize_spmv_coo_op(op)]


def _spmv_coo_op_transposes(op):
    return _normalize_spmv_coo_op(op) in (SPMV_COO_OP_TRANS, SPMV_COO_OP_CONJ_TRANS)


def _normalize_spmv_coo_index_fallback_policy(index_fallback_policy):
    policy = str(index_fallback_policy).lower()
    if policy not in ("auto", "strict"):
        raise ValueError("index_fallback_policy must be 'auto' or 'strict'")
    return policy


class PreparedCoo:
    """Sorted COO + row-run bounds ``seg_starts`` (spmv_coo.py:77-129), plus
    the row pointer of the sorted entries for the CSR kernels."""

    __slots__ = ("data", "row", "col", "shape", "n_rows", "n_cols", "nnz", "seg_starts",
                 "n_segs", "use_seg_kernel", "op", "transpose", "index_fallback_policy",
                 "index_fallback_applied", "index_fallback_reason", "_csr")

    def __init__(self, data, row, col, shape, seg_starts=None, transpose=False, op=None,
                 index_fallback_policy="auto", index_fallback_applied=False,
                 index_fallback_reason=None):
        self.data = data
        self.row = row
        self.col = col
        self.shape = (int(shape[0]), int(shape[1]))
        self.n_rows, self.n_cols = self.shape
        self.nnz = int(data.numel())
        self.seg_starts = seg_starts
        if seg_starts is None:
            self.n_segs = 0
            self.use_seg_kernel = False
        else:
            self.n_segs = int(seg_starts.numel()) - 1
            self.use_seg_kernel = self.n_segs > 0
        self.op = _normalize_spmv_coo_op(op, transpose=transpose)
        self.transpose = _spmv_coo_op_transposes(self.op)
        self.index_fallback_policy = str(index_fallback_policy).lower()
        self.index_fallback_applied = bool(index_fallback_applied)
        self.index_fallback_reason = index_fallback_reason
        self._csr = None


def _sort_coo_lex(data, row, col, n_cols):
    """Entries ordered by (row, col); stable, so repeated pairs keep their
    input order (the reference's argsort is unstable, spmv_coo.py:297-309)."""
    index_dtype = torch.promote_types(row.dtype, col.dtype)
    if data.numel() == 0:
        return data.contiguous(), row.to(index_dtype).contiguous(), col.to(index_dtype).contiguous()
    key = row.to(torch.int64) * max(1, int(n_cols)) + col.to(torch.int64)
    order = torch.argsort(key, stable=True)
    return (data[order].contiguous(), row[order].to(index_dtype).contiguous(),
            col[order].to(index_dtype).contiguous())


def _seg_starts_from_sorted_rows(row, nnz, device):
    """Bounds of constant-row runs of sorted COO (spmv_coo.py:312-327)."""
    if nnz == 0:
        return None
    index_dtype = row.dtype if nnz <= _INDEX_LIMIT_INT32 else torch.int64
    breaks = torch.nonzero(row[1:] != row[:-1], as_tuple=False).flatten().to(index_dtype) + 1
    return torch.cat([torch.zeros(1, dtype=index_dtype, device=device), breaks,
                      torch.tensor([nnz], dtype=index_dtype, device=device)])


def _prepare_coo_tensors(data, row, col, shape, sort_by_row):
    """Validation of spmv_coo.py:330-362, same order and messages."""
    if not all(torch.is_tensor(t) for t in (data, row, col)):
        raise TypeError("data, row, col must all be torch.Tensor")
    if data.ndim != 1 or row.ndim != 1 or col.ndim != 1:
        raise ValueError("data, row, col must be 1D")
    if not all(t.is_cuda for t in (data, row, col)):
        raise ValueError("data, row, col must be CUDA tensors")
    n_rows, n_cols = int(shape[0]), int(shape[1])
    if data.numel() != row.numel() or data.numel() != col.numel():
        raise ValueError("data, row, col must have the same length")
    if data.dtype not in SUPPORTED_SPMV_COO_VALUE_DTYPES:
        raise TypeError(_spmv_coo_dtype_error_message())
    if row.dtype not in SUPPORTED_INDEX_DTYPES:
        raise TypeError("row dtype must be torch.int32 or torch.int64")
    if col.dtype not in SUPPORTED_INDEX_DTYPES:
        raise TypeError("col dtype must be torch.int32 or torch.int64")
    # both orders are sorted here (deterministic sums); seg_starts only when asked
    data, kr, kc = _sort_coo_lex(data, row, col, n_cols)
    seg = _seg_starts_from_sorted_rows(kr, data.numel(), data.device) if sort_by_row else None
    if data.numel() > 0:
        r64, c64 = kr.to(torch.int64), kc.to(torch.int64)
        if int(r64.min().item()) < 0 or int(r64.max().item()) >= n_rows:
            raise IndexError("row indices out of range")
        if int(c64.min().item()) < 0 or int(c64.max().item()) >= n_cols:
            raise IndexError("col indices out of range")
    return data, kr, kc, seg


def prepare_spmv_coo(data, row, col, shape, sort_by_row=True, transpose=False, op=None,
                     index_fallback_policy="auto"):
    """Sorted COO (+ row runs) for repeated SpMV (spmv_coo.py:365-402)."""
    index_fallback_policy = _normalize_spmv_coo_index_fallback_policy(index_fallback_policy)
    op_code = _normalize_spmv_coo_op(op, transpose=transpose)
    if op is not None and bool(transpose) and op_code == SPMV_COO_OP_NON:
        raise ValueError("transpose=True conflicts with op=non")
    transpose = _spmv_coo_op_transposes(op_code)
    if transpose:
        shape = (int(shape[1]), int(shape[0]))
        row, col = col, row
        if op_code == SPMV_COO_OP_CONJ_TRANS and _is_complex_dtype(data.dtype):
            data = data.conj()
            if hasattr(data, "resolve_conj"):
                data = data.resolve_conj()
    d, kr, kc, seg = _prepare_coo_tensors(data, row, col, shape, sort_by_row)
    return PreparedCoo(d, kr, kc, shape, seg_starts=seg, transpose=transpose, op=op_code,
                       index_fallback_policy=index_fallback_policy)


def _validate_x_coo(x, prepared):
    """x checks of spmv_coo.py:405-420."""
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


def _csr_of(prepared):
    """PreparedCsrSpmv over the sorted entries (row pointer from the rows)."""
    if prepared._csr is None:
        from .spmv_csr import prepare_spmv_csr
        counts = torch.bincount(prepared.row.to(torch.int64), minlength=prepared.n_rows)
        ptr_dtype = torch.int32 if prepared.nnz <= _INDEX_LIMIT_INT32 else torch.int64
        indptr = torch.zeros(prepared.n_rows + 1, dtype=torch.int64, device=prepared.data.device)
        torch.cumsum(counts, 0, out=indptr[1:])
        prepared._csr = prepare_spmv_csr(prepared.data, prepared.col, indptr.to(ptr_dtype),
                                         prepared.shape)
    return prepared._csr


def _hip_spmv_coo_kernel(prepared, x, block_size=256, num_warps=4, block_inner=128):
    """y = A x over the sorted entries: the CSR SpMV with native accumulation
    (the reference's COO kernels accumulate in the value dtype)."""
    from .spmv_csr import _hip_spmv_csr_impl_opt_prepared
    if prepared.n_rows == 0:
        return torch.empty(0, dtype=prepared.data.dtype, device=prepared.data.device)
    if prepared.nnz == 0:
        return torch.zeros(prepared.n_rows, dtype=prepared.data.dtype, device=prepared.data.device)
    return _hip_spmv_csr_impl_opt_prepared(_csr_of(prepared), x)


_triton_spmv_coo_kernel = _hip_spmv_coo_kernel  # the reference's launcher name


def _spmv_coo_uses_int64_indices(prepared):
    return (prepared.row.dtype == torch.int64 or prepared.col.dtype == torch.int64
            or (prepared.seg_starts is not None and prepared.seg_starts.dtype == torch.int64))


def _spmv_coo_int32_fallback_blocker(prepared):
    """Why int32 indices cannot stand in (spmv_coo.py:506-522), or None."""
    for name, tensor in (("row", prepared.row), ("col", prepared.col)):
        if tensor.numel() > 0:
            lo, hi = int(tensor.min().item()), int(tensor.max().item())
            if lo < 0 or hi > _INDEX_LIMIT_INT32:
                return (f"{name} index range [{lo}, {hi}] cannot fit int32 "
                        f"for shape={prepared.shape}")
    if prepared.seg_starts is not None and prepared.seg_starts.numel() > 0:
        max_offset = int(prepared.seg_starts[-1].item())
        if max_offset > _INDEX_LIMIT_INT32:
            return f"COO nnz offset {max_offset} cannot fit int32 for shape={prepared.shape}"
    if prepared.n_rows > _INDEX_LIMIT_INT32 or prepared.n_cols > _INDEX_LIMIT_INT32:
        return f"shape {prepared.shape} cannot fit int32 row/col metadata"
    return None


def _spmv_coo_prepared_with_int32_indices(prepared, reason):
    blocker = _spmv_coo_int32_fallback_blocker(prepared)
    if blocker is not None:
        raise RuntimeError(f"native int64 COO SpMV failed and int32 fallback is unsafe: {blocker}")
    seg = None if prepared.seg_starts is None else prepared.seg_starts.to(torch.int32).contiguous()
    return PreparedCoo(data=prepared.data, row=prepared.row.to(torch.int32).contiguous(),
                       col=prepared.col.to(torch.int32).contiguous(), shape=prepared.shape,
                       seg_starts=seg, transpose=prepared.transpose, op=prepared.op,
                       index_fallback_policy=prepared.index_fallback_policy,
                       index_fallback_applied=True, index_fallback_reason=str(reason))


def _run_spmv_coo_prepared_with_fallback(prepared, x, block_size, num_warps, block_inner):
    try:
        return _triton_spmv_coo_kernel(prepared, x, block_size=block_size, num_warps=num_warps,
                                       block_inner=block_inner)
    except Exception as exc:
        if prepared.index_fallback_policy != "auto" or not _spmv_coo_uses_int64_indices(prepared):
            raise
        fb = _spmv_coo_prepared_with_int32_indices(prepared, exc)
        return _triton_spmv_coo_kernel(fb, x, block_size=block_size, num_warps=num_warps,
                                       block_inner=block_inner)


def flagsparse_spmv_coo(data=None, row=None, col=None, x=None, shape=None, out=None,
                        return_time=False, prepared=None, sort_by_row=True, block_size=256,
                        num_warps=4, block_inner=128, transpose=None, op=None,
                        index_fallback_policy="auto"):
    """y = op(A) x for COO A (spmv_coo.py:575-680).  ``block_size`` /
    ``num_warps`` / ``block_inner`` are validated as in the reference and do
    not change the HIP schedule."""
    op_explicit = op is not None
    op_code = _normalize_spmv_coo_op(op, transpose=False if transpose is None else bool(transpose))
    if op_explicit and transpose is not None and bool(transpose) != _spmv_coo_op_transposes(op_code):
        raise ValueError("transpose conflicts with op")
    if prepared is None:
        if any(a is None for a in (data, row, col, x, shape)):
            raise ValueError("data, row, col, x, shape required when prepared is None")
        prepared = prepare_spmv_coo(data, row, col, shape, sort_by_row=sort_by_row, op=op_code,
                                    index_fallback_policy=index_fallback_policy)
    else:
        if x is None:
            raise TypeError("x is required when prepared is set")
        if op_explicit and op_code != prepared.op:
            raise ValueError(f"op={_spmv_coo_op_to_name(op_code)} does not match "
                             f"prepared.op={_spmv_coo_op_to_name(prepared.op)}")
        if not op_explicit and transpose is not None and bool(transpose) != prepared.transpose:
            raise ValueError(f"transpose={bool(transpose)} does not match "
                             f"prepared.transpose={prepared.transpose}")
        if shape is None:
            shape = prepared.shape
        sh = (int(shape[0]), int(shape[1]))
        if sh != prepared.shape:
            raise ValueError(f"shape {sh} does not match prepared.shape {prepared.shape}")
    x = _validate_x_coo(x, prepared)
    if num_warps not in (1, 2, 4, 8, 16, 32):
        raise ValueError("num_warps must be a power of 2 in [1, 32]")
    if block_inner <= 0 or (block_inner & (block_inner - 1)) != 0:
        raise ValueError("block_inner must be a positive power of 2")
    t0 = None
    if return_time:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
    y = _run_spmv_coo_prepared_with_fallback(prepared, x, block_size=block_size,
                                             num_warps=num_warps, block_inner=block_inner)
    elapsed_ms = None
    if return_time:
        torch.cuda.synchronize()
        elapsed_ms = (time.perf_counter() - t0) * 1000.0
    if out is not None:
        if out.shape != y.shape or out.dtype != y.dtype:
            raise ValueError("out shape/dtype must match result")
        out.copy_(y)
        y = out
    if return_time:
        return y, elapsed_ms
    return y


__all__ = ["SUPPORTED_SPMV_COO_VALUE_DTYPES", "PreparedCoo", "prepare_spmv_coo",
           "flagsparse_spmv_coo"]
