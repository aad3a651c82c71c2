"""rocSPARSE baseline (libfs_vendor.so), timed beside this library's kernels.

The reference's harness times cuSPARSE through CuPy next to its own kernels
(benchmarks.py:345-507, spmm_csr.py:2351-2564: the ``cusparse_*`` fields).
On MI355X the vendor library is rocSPARSE; ``VendorSpmv`` / ``VendorSpmm``
bind one matrix and its dense operands once (buffer-size and preprocess
stages) and ``run()`` enqueues only the compute stage on torch's current
stream.  This is a baseline for measurements, not a path of the operators.
"""

from __future__ import annotations

import ctypes
import os

import torch

from . import _lib

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libfs_vendor.so")

# rocsparse_spmv_alg / rocsparse_spmm_alg values
SPMV_ALGS = {"adaptive": 2, "rowsplit": 3, "lrb": 7, "nnzsplit": 8}
SPMM_ALGS = {"default": 0, "csr": 1, "row_split": 5, "nnz_split": 6, "merge_path": 7}

_lib_handle = None
_load_error = None


def load():
    """The loaded library, or None with the reason in ``unavailable_reason()``."""
    global _lib_handle, _load_error
    if _lib_handle is not None or _load_error is not None:
        return _lib_handle
    if not os.path.exists(LIB_PATH):
        _load_error = f"{LIB_PATH} is not built"
        return None
    try:
        lib = ctypes.CDLL(LIB_PATH)
    except OSError as exc:
        _load_error = str(exc)
        return None
    vp = ctypes.c_void_p
    lib.fsv_spmv_setup.argtypes = [ctypes.POINTER(_lib.FsCsr), vp, vp, ctypes.c_int, vp,
                                   ctypes.POINTER(vp)]
    lib.fsv_spmm_setup.argtypes = [ctypes.POINTER(_lib.FsCsr), vp, ctypes.c_int64, vp,
                                   ctypes.c_int, vp, ctypes.POINTER(vp)]
    lib.fsv_sddmm_setup.argtypes = [ctypes.POINTER(_lib.FsCsr), vp, vp, ctypes.c_int64,
                                    ctypes.c_int, vp, ctypes.POINTER(vp)]
    lib.fsv_run.argtypes = [vp, ctypes.c_double, ctypes.c_double]
    lib.fsv_destroy.argtypes = [vp]
    lib.fsv_destroy.restype = None
    lib.fsv_last_error.restype = ctypes.c_char_p
    for f in ("fsv_spmv_setup", "fsv_spmm_setup", "fsv_sddmm_setup", "fsv_run"):
        getattr(lib, f).restype = ctypes.c_int
    _lib_handle = lib
    return lib


def unavailable_reason():
    load()
    return _load_error


def _check(lib, rc, what):
    if rc != 0:
        raise RuntimeError(f"rocSPARSE {what}: {lib.fsv_last_error().decode()}")


class _Bound:
    def __init__(self):
        self._ctx = None
        self._keep = ()

    def run(self, alpha=1.0, beta=0.0):
        _check(self._lib, self._lib.fsv_run(self._ctx, float(alpha), float(beta)), "compute")

    def __del__(self):
        if self._ctx is not None and self._lib is not None:
            self._lib.fsv_destroy(self._ctx)
            self._ctx = None


class VendorSpmv(_Bound):
    """y = A x with rocSPARSE's CSR SpMV (algorithm by name, see SPMV_ALGS)."""

    def __init__(self, data, indices, indptr, shape, x, y, alg="adaptive"):
        super().__init__()
        self._lib = load()
        if self._lib is None:
            raise RuntimeError(f"rocSPARSE baseline unavailable: {_load_error}")
        csr = _lib.make_csr(data, indices, indptr, int(shape[0]), int(shape[1]))
        ctx = ctypes.c_void_p()
        rc = self._lib.fsv_spmv_setup(ctypes.byref(csr), x.data_ptr(), y.data_ptr(),
                                      SPMV_ALGS[alg], _lib.stream_handle(x.device),
                                      ctypes.byref(ctx))
        _check(self._lib, rc, f"spmv setup ({alg})")
        self._ctx = ctx
        self._keep = (data, indices, indptr, x, y, csr)


class VendorSpmm(_Bound):
    """C = A B with rocSPARSE's CSR SpMM (row-major B and C)."""

    def __init__(self, data, indices, indptr, shape, B, C, alg="default"):
        super().__init__()
        self._lib = load()
        if self._lib is None:
            raise RuntimeError(f"rocSPARSE baseline unavailable: {_load_error}")
        if not (B.is_contiguous() and C.is_contiguous()):
            raise ValueError("B and C must be row-major contiguous")
        csr = _lib.make_csr(data, indices, indptr, int(shape[0]), int(shape[1]))
        ctx = ctypes.c_void_p()
        rc = self._lib.fsv_spmm_setup(ctypes.byref(csr), B.data_ptr(), int(B.shape[1]),
                                      C.data_ptr(), SPMM_ALGS[alg],
                                      _lib.stream_handle(B.device), ctypes.byref(ctx))
        _check(self._lib, rc, f"spmm setup ({alg})")
        self._ctx = ctx
        self._keep = (data, indices, indptr, B, C, csr)


class VendorSddmm(_Bound):
    """out[p] = alpha <x[row p], y[col p]> + beta out[p] with rocSPARSE's
    SDDMM over the CSR pattern (indices, indptr); x [n_rows, k] and y
    [n_cols, k] row-major; `out` holds the pattern's values."""

    def __init__(self, indices, indptr, shape, x, y, out):
        super().__init__()
        self._lib = load()
        if self._lib is None:
            raise RuntimeError(f"rocSPARSE baseline unavailable: {_load_error}")
        if not (x.is_contiguous() and y.is_contiguous() and out.is_contiguous()):
            raise ValueError("x, y and out must be contiguous")
        csr = _lib.make_csr(out, indices, indptr, int(shape[0]), int(shape[1]))
        ctx = ctypes.c_void_p()
        rc = self._lib.fsv_sddmm_setup(ctypes.byref(csr), x.data_ptr(), y.data_ptr(),
                                       int(x.shape[1]), 0, _lib.stream_handle(x.device),
                                       ctypes.byref(ctx))
        _check(self._lib, rc, "sddmm setup")
        self._ctx = ctx
        self._keep = (indices, indptr, x, y, out, csr)


def available() -> bool:
    return load() is not None and torch.cuda.is_available()


def spmv_baseline(data, indices, indptr, shape, x, op_fn, alg="adaptive"):
    """(values, callable or None, reason) for the harnesses: rocSPARSE y = A x,
    or the reason it is unavailable for this dtype."""
    if data.dtype not in (torch.float32, torch.float64, torch.complex64, torch.complex128):
        return None, None, "float16/bfloat16 are not timed against rocSPARSE"
    if load() is None:
        return None, None, f"rocSPARSE baseline unavailable: {_load_error}"
    y = torch.empty(int(shape[0]), dtype=data.dtype, device=data.device)
    try:
        bound = VendorSpmv(data, indices, indptr, shape, x, y, alg)
    except Exception as exc:  # setup errors are reported, not raised, by the harness
        return None, None, str(exc)

    def run():
        bound.run()
        return op_fn(y)
    return run(), run, None


def spmm_baseline(data, indices, indptr, shape, B, alg="default"):
    """(values, callable or None, reason): rocSPARSE C = A B for row-major B."""
    if data.dtype not in (torch.float32, torch.float64, torch.complex64, torch.complex128):
        return None, None, "float16/bfloat16 are not timed against rocSPARSE"
    if load() is None:
        return None, None, f"rocSPARSE baseline unavailable: {_load_error}"
    Bc = B.contiguous()
    C = torch.empty((int(shape[0]), int(Bc.shape[1])), dtype=data.dtype, device=data.device)
    try:
        bound = VendorSpmm(data, indices, indptr, shape, Bc, C, alg)
    except Exception as exc:
        return None, None, str(exc)

    def run():
        bound.run()
        return C
    return run(), run, None
