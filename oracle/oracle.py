"""CPU oracle for the CSR SpMV / SpMM hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker (or the timed CPU
baseline).  The product package never imports it.

It wraps ``csr_oracle.c`` (built into ``liboracle.so`` by ``oracle/Makefile``)
with numpy widening: values become float64 (complex -> interleaved float64
pairs), indices become int64.  Each function names the reference behaviour it
restates (paths relative to the reference repository).
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)
_f32p = ctypes.POINTER(ctypes.c_float)
_u16p = ctypes.POINTER(ctypes.c_uint16)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (gcc)."""
    if force or not os.path.exists(_LIB_PATH):
        cmd = ["make", "-s", "-C", _HERE] + (["-B"] if force else [])
        subprocess.run(cmd, check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        i64, f64, vp = ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
        sig = {
            "orc_spmv_f64": (None, [i64, vp, vp, vp, vp, vp]),
            "orc_spmv_abs_f64": (None, [i64, vp, vp, vp, vp, vp]),
            "orc_spmv_c128": (None, [i64, vp, vp, vp, vp, vp]),
            "orc_spmm_f64": (None, [i64, i64, vp, vp, vp, vp, i64, vp, i64]),
            "orc_spmm_abs_f64": (None, [i64, i64, vp, vp, vp, vp, i64, vp, i64]),
            "orc_spmm_c128": (None, [i64, i64, vp, vp, vp, vp, i64, vp, i64]),
            "orc_sddmm_f64": (None, [i64, i64, vp, vp, vp, vp, vp, f64, f64, vp, vp]),
            "orc_csr_transpose": (None, [i64, i64, vp, vp, vp, vp, vp]),
            "orc_spmv_opt_bucket": (ctypes.c_int, [i64, ctypes.c_int]),
            "orc_spmm_opt_bucket": (ctypes.c_int, [i64]),
            "orc_spmm_alg2_bucket": (ctypes.c_int, [i64]),
            "orc_split_parts": (i64, [i64, vp, vp, i64, i64, vp, vp, vp, vp]),
            "orc_row_features": (None, [i64, i64, vp, vp]),
            "orc_mlp_forward": (None, [i64, i64, i64, i64] + [vp] * 10),
            "orc_mlp_backward": (f64, [i64, i64, i64, i64] + [vp] * 14),
            "orc_adam_step": (None, [i64, vp, vp, vp, vp, f64, f64, f64, f64, i64]),
            "orc_bf16_split": (None, [i64, vp, vp, vp]),
            "orc_bf16_join": (None, [i64, vp, vp, vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _i64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a), dtype=np.int64)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a), dtype=np.float64)


def _c128_ri(a) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a), dtype=np.complex128)
    return a.view(np.float64)


def is_complex(a) -> bool:
    return np.iscomplexobj(np.asarray(a))


# ----------------------------------------------------------------- SpMV / SpMM

def spmv(indptr, indices, data, x) -> np.ndarray:
    """y = A x in fp64 (complex128 for complex data), sequential CSR order.

    Restates spmv_csr.py:158-184 (real) and :187-219 (complex); the default
    path's fp64 widening of fp32 inputs is spmv_csr.py:116-122, 664-689.
    """
    indptr, indices = _i64(indptr), _i64(indices)
    n = indptr.size - 1
    if is_complex(data) or is_complex(x):
        d, xx = _c128_ri(data), _c128_ri(x)
        y = np.zeros(2 * n, dtype=np.float64)
        lib().orc_spmv_c128(n, _p(indptr), _p(indices), _p(d), _p(xx), _p(y))
        return y.view(np.complex128)
    d, xx = _f64(data), _f64(x)
    y = np.zeros(n, dtype=np.float64)
    lib().orc_spmv_f64(n, _p(indptr), _p(indices), _p(d), _p(xx), _p(y))
    return y


def spmv_abs(indptr, indices, data, x) -> np.ndarray:
    """Per-row condition denominator sum_j |a_ij x_j| (|.| of complex products)."""
    indptr, indices = _i64(indptr), _i64(indices)
    n = indptr.size - 1
    if is_complex(data) or is_complex(x):
        rows = np.repeat(np.arange(n), np.diff(indptr))
        prod = np.abs(np.asarray(data, np.complex128) * np.asarray(x, np.complex128)[indices])
        return np.bincount(rows, weights=prod, minlength=n).astype(np.float64)
    d, xx = _f64(data), _f64(x)
    s = np.zeros(n, dtype=np.float64)
    lib().orc_spmv_abs_f64(n, _p(indptr), _p(indices), _p(d), _p(xx), _p(s))
    return s


def spmm(indptr, indices, data, B) -> np.ndarray:
    """C = A B in fp64 (complex128 for complex), row-major (spmm_csr.py:92-134)."""
    indptr, indices = _i64(indptr), _i64(indices)
    n = indptr.size - 1
    B = np.asarray(B)
    nd = B.shape[1]
    if is_complex(data) or is_complex(B):
        d = _c128_ri(data)
        Bri = np.ascontiguousarray(B, dtype=np.complex128)
        C = np.zeros((n, nd), dtype=np.complex128)
        lib().orc_spmm_c128(n, nd, _p(indptr), _p(indices), _p(d),
                            _p(Bri.view(np.float64)), nd, _p(C.view(np.float64)), nd)
        return C
    d = _f64(data)
    Bf = np.ascontiguousarray(B, dtype=np.float64)
    C = np.zeros((n, nd), dtype=np.float64)
    lib().orc_spmm_f64(n, nd, _p(indptr), _p(indices), _p(d), _p(Bf), nd, _p(C), nd)
    return C


def sddmm(indptr, indices, x, y, data=None, alpha=1.0, beta=0.0):
    """(out, abs_out) of SDDMM in fp64 (sddmm_csr.py:108-151, 423-437)."""
    indptr, indices = _i64(indptr), _i64(indices)
    n = indptr.size - 1
    xf = np.ascontiguousarray(x, dtype=np.float64)
    yf = np.ascontiguousarray(y, dtype=np.float64)
    k = xf.shape[1]
    nnz = indices.size
    out = np.zeros(nnz)
    ab = np.zeros(nnz)
    d = _f64(data) if data is not None else None
    lib().orc_sddmm_f64(n, k, _p(indptr), _p(indices), _p(xf), _p(yf),
                        _p(d) if d is not None else None, float(alpha), float(beta), _p(out),
                        _p(ab))
    return out, ab


def spmm_abs(indptr, indices, data, B) -> np.ndarray:
    indptr, indices = _i64(indptr), _i64(indices)
    n = indptr.size - 1
    B = np.asarray(B)
    nd = B.shape[1]
    if is_complex(data) or is_complex(B):
        rows = np.repeat(np.arange(n), np.diff(indptr))
        S = np.zeros((n, nd))
        np.add.at(S, rows, np.abs(np.asarray(data, np.complex128)[:, None] *
                                  np.asarray(B, np.complex128)[indices]))
        return S
    d = _f64(data)
    Bf = np.ascontiguousarray(B, dtype=np.float64)
    S = np.zeros((n, nd), dtype=np.float64)
    lib().orc_spmm_abs_f64(n, nd, _p(indptr), _p(indices), _p(d), _p(Bf), nd, _p(S), nd)
    return S


# ------------------------------------------------------------------ transpose

def csr_transpose(indptr, indices, shape):
    """CSR of A^T with stable per-column order; returns (t_indptr, t_indices, perm).

    Restates _transpose_csr_for_spmv (spmv_csr.py:456-494): entries sorted by
    column with a stable sort, so data_T = data[perm].
    """
    indptr, indices = _i64(indptr), _i64(indices)
    n_rows, n_cols = int(shape[0]), int(shape[1])
    nnz = int(indptr[-1]) if indptr.size else 0
    t_indptr = np.zeros(n_cols + 1, dtype=np.int64)
    t_indices = np.zeros(max(nnz, 1), dtype=np.int64)
    perm = np.zeros(max(nnz, 1), dtype=np.int64)
    lib().orc_csr_transpose(n_rows, n_cols, _p(indptr), _p(indices), _p(t_indptr),
                            _p(t_indices), _p(perm))
    return t_indptr, t_indices[:nnz], perm[:nnz]


# -------------------------------------------------------------------- buckets

def spmv_opt_bucket(lengths, fp64=False) -> np.ndarray:
    """Bucket id per row for the SpMV opt path (spmv_csr.py:128-153, 345-391)."""
    L = lib()
    return np.array([L.orc_spmv_opt_bucket(int(v), int(bool(fp64))) for v in lengths],
                    dtype=np.int64)


def spmm_opt_bucket(lengths) -> np.ndarray:
    """Bucket id per row for the SpMM opt path (spmm_csr.py:395-401, 444-477)."""
    L = lib()
    return np.array([L.orc_spmm_opt_bucket(int(v)) for v in lengths], dtype=np.int64)


def spmm_alg2_bucket(lengths) -> np.ndarray:
    """Bucket id per row for alg2 (spmm_csr_opt_alg2.py:60-166, 307-335)."""
    L = lib()
    return np.array([L.orc_spmm_alg2_bucket(int(v)) for v in lengths], dtype=np.int64)


def split_parts(indptr, long_rows, part_nnz=256):
    """(part_rows, part_starts, part_ends, row_part_ptr) (spmm_csr.py:417-441)."""
    indptr, long_rows = _i64(indptr), _i64(long_rows)
    lengths = np.diff(indptr)[long_rows] if long_rows.size else np.zeros(0, np.int64)
    total = int(np.sum((lengths + part_nnz - 1) // part_nnz)) if long_rows.size else 0
    cap = max(total, 1)
    pr = np.zeros(cap, np.int64)
    ps = np.zeros(cap, np.int64)
    pe = np.zeros(cap, np.int64)
    rpp = np.zeros(long_rows.size + 1, np.int64)
    n = lib().orc_split_parts(long_rows.size, _p(long_rows), _p(indptr), int(part_nnz),
                              cap, _p(pr), _p(ps), _p(pe), _p(rpp))
    return pr[:n], ps[:n], pe[:n], rpp


# ------------------------------------------------------------------- features

def row_features(indptr, n_cols) -> np.ndarray:
    """The 16 row-length statistics the kernel selector reads (see csr_oracle.c)."""
    indptr = _i64(indptr)
    f = np.zeros(16, dtype=np.float64)
    lib().orc_row_features(indptr.size - 1, int(n_cols), _p(indptr), _p(f))
    return f


# ------------------------------------------------------------------------ MLP

def mlp_forward(X, W1, b1, W2, b2, W3, b3):
    X = _f64(X)
    batch, ind = X.shape
    hid = W1.shape[1]
    outd = W3.shape[1]
    H1 = np.zeros((batch, hid))
    H2 = np.zeros((batch, hid))
    Y = np.zeros((batch, outd))
    args = [_f64(a) for a in (W1, b1, W2, b2, W3, b3)]
    lib().orc_mlp_forward(batch, ind, hid, outd, _p(X), *[_p(a) for a in args],
                          _p(H1), _p(H2), _p(Y))
    return H1, H2, Y


def mlp_backward(X, W2, W3, H1, H2, Y, T, M):
    X = _f64(X)
    batch, ind = X.shape
    hid = W2.shape[0]
    outd = W3.shape[1]
    g = [np.zeros((ind, hid)), np.zeros(hid), np.zeros((hid, hid)), np.zeros(hid),
         np.zeros((hid, outd)), np.zeros(outd)]
    ins = [_f64(a) for a in (W2, W3, H1, H2, Y, T, M)]
    loss = lib().orc_mlp_backward(batch, ind, hid, outd, _p(X), *[_p(a) for a in ins],
                                  *[_p(a) for a in g])
    return loss, g


def adam_step(p, g, m, v, lr, beta1, beta2, eps, step):
    """In-place Adam update of float64 arrays p, m, v."""
    lib().orc_adam_step(p.size, _p(p), _p(_f64(g)), _p(m), _p(v), lr, beta1, beta2,
                        eps, int(step))


# -------------------------------------------------------- split value storage

def bf16_split(values):
    """hi = bf16(v), lo = bf16(v - hi) as uint16 bit patterns (round to nearest even)."""
    v = np.ascontiguousarray(values, dtype=np.float32)
    hi = np.zeros(v.size, np.uint16)
    lo = np.zeros(v.size, np.uint16)
    lib().orc_bf16_split(v.size, _p(v), _p(hi), _p(lo))
    return hi, lo


def bf16_join(hi, lo=None) -> np.ndarray:
    hi = np.ascontiguousarray(hi, dtype=np.uint16)
    out = np.zeros(hi.size, np.float64)
    lo_p = _p(np.ascontiguousarray(lo, dtype=np.uint16)) if lo is not None else None
    lib().orc_bf16_join(hi.size, _p(hi), lo_p, _p(out))
    return out


# --------------------------------------------------------------- error model

def allclose_ratio(actual, reference, atol, rtol) -> float:
    """max(|d| / (atol + rtol |ref|)) -- tests/test_spmv.py:179-184."""
    a = np.asarray(actual)
    r = np.asarray(reference)
    if a.size == 0:
        return 0.0
    d = np.abs(a.astype(np.complex128 if np.iscomplexobj(a) else np.float64) - r)
    return float(np.max(d / (atol + rtol * np.abs(r))))


def rel_err_per_row(actual, ref64, cond):
    """Per-row error |y - y64| / max(sum|a x|, tiny) (condition normalised)."""
    a = np.asarray(actual).astype(np.complex128 if np.iscomplexobj(actual) else np.float64)
    d = np.abs(a - ref64)
    return d / np.maximum(cond, np.finfo(np.float64).tiny)
