"""Writes tests/golden/csr_golden.npz: small CSR cases with exactly rounded
expected outputs.

The expected values are computed with exact rational arithmetic
(fractions.Fraction over the stored binary values), then rounded once to
fp64 (``y_exact``) and once to fp32 (``y_f32``).  They do not come from the
oracle, so they pin it.  Cases follow the reference's own known-answer and
edge cases (paths relative to the reference repository):

* ``ka2x2``    -- A = [[1, 2], [0, 3]] as data [1,2,3], indices [0,1,1],
                  indptr [0,2,3] (tests/test_spmm.py:720-723);
* ``longrow``  -- row 0 = columns 0..1535, row 1 = {7, 129, 511, 1024},
                  shape (2, 2048) (tests/test_spmm.py:401-413);
* ``unsorted_dup`` -- unsorted and duplicate columns (the reference's random
                  builder makes them, _common.py:223-238);
* ``empty_rows``   -- interleaved empty rows;
* ``all_empty``    -- nnz = 0 (tests/test_spmm.py:773-790);
* ``cancel``       -- rows whose exact sum cancels to (near) zero;
* ``complex``      -- complex128 values (spmv_csr.py:187-219).

Run: python tests/golden/make_golden.py  (deterministic, seed 20261202)
"""

from fractions import Fraction
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def exact_spmv(indptr, indices, data, x):
    n = len(indptr) - 1
    out = []
    for r in range(n):
        s = Fraction(0)
        for j in range(indptr[r], indptr[r + 1]):
            s += Fraction(float(data[j])) * Fraction(float(x[indices[j]]))
        out.append(s)
    return out


def exact_spmv_complex(indptr, indices, data, x):
    n = len(indptr) - 1
    re, im = [], []
    for r in range(n):
        sr, si = Fraction(0), Fraction(0)
        for j in range(indptr[r], indptr[r + 1]):
            ar, ai = Fraction(float(data[j].real)), Fraction(float(data[j].imag))
            xr, xi = Fraction(float(x[indices[j]].real)), Fraction(float(x[indices[j]].imag))
            sr += ar * xr - ai * xi
            si += ar * xi + ai * xr
        re.append(sr)
        im.append(si)
    return re, im


def exact_spmm(indptr, indices, data, B):
    n = len(indptr) - 1
    nd = B.shape[1]
    out = np.empty((n, nd), dtype=object)
    for r in range(n):
        for k in range(nd):
            s = Fraction(0)
            for j in range(indptr[r], indptr[r + 1]):
                s += Fraction(float(data[j])) * Fraction(float(B[indices[j], k]))
            out[r, k] = s
    return out


def to_f64(fr):
    return np.array([float(v) for v in np.ravel(fr)], dtype=np.float64).reshape(np.shape(fr))


def to_f32(fr):
    # float(Fraction) is correctly rounded to fp64; rounding that again to fp32
    # could double-round, so round the rational directly.
    out = []
    for v in np.ravel(fr):
        f64 = float(v)
        c = np.float32(f64)
        # pick the fp32 neighbour closest to the exact rational
        cands = {c, np.nextafter(c, np.float32(np.inf)), np.nextafter(c, np.float32(-np.inf))}
        best = min(cands, key=lambda t: (abs(Fraction(float(t)) - v), int(np.float32(t).view(np.uint32)) & 1))
        out.append(best)
    return np.array(out, dtype=np.float32).reshape(np.shape(fr))


def main():
    rng = np.random.Generator(np.random.PCG64(20261202))
    cases = {}

    def add(name, indptr, indices, data, shape, x=None, B=None):
        indptr = np.asarray(indptr, np.int64)
        indices = np.asarray(indices, np.int64)
        c = {"indptr": indptr, "indices": indices, "data": data,
             "shape": np.asarray(shape, np.int64)}
        if x is None:
            x = rng.uniform(-1, 1, shape[1]).astype(data.dtype)
        c["x"] = x
        if np.iscomplexobj(data):
            re, im = exact_spmv_complex(indptr, indices, data, x)
            c["y_exact"] = to_f64(re) + 1j * to_f64(im)
        else:
            ex = exact_spmv(indptr, indices, data, x)
            c["y_exact"] = to_f64(ex)
            c["y_f32"] = to_f32(ex)
            if B is None:
                B = rng.uniform(-1, 1, (shape[1], 5)).astype(data.dtype)
            c["B"] = B
            exm = exact_spmm(indptr, indices, data, B)
            c["C_exact"] = to_f64(exm)
            c["C_f32"] = to_f32(exm)
        for k, v in c.items():
            cases[f"{name}__{k}"] = v

    add("ka2x2", [0, 2, 3], [0, 1, 1], np.array([1.0, 2.0, 3.0], np.float32), (2, 2),
        x=np.array([0.5, -1.25], np.float32))

    long_cols = np.concatenate([np.arange(1536), [7, 129, 511, 1024]])
    add("longrow", [0, 1536, 1540], long_cols,
        rng.uniform(-1, 1, long_cols.size).astype(np.float32), (2, 2048))

    lens = np.array([3, 5, 1, 4])
    ip = np.concatenate([[0], np.cumsum(lens)])
    cols = np.array([5, 1, 5, 0, 7, 7, 2, 2, 3, 6, 6, 0, 1])
    add("unsorted_dup", ip, cols, rng.uniform(-1, 1, cols.size).astype(np.float32), (4, 8))

    lens = np.array([0, 2, 0, 0, 3, 0, 1, 0])
    ip = np.concatenate([[0], np.cumsum(lens)])
    cols = np.array([1, 4, 0, 2, 5, 3])
    add("empty_rows", ip, cols, rng.uniform(-1, 1, cols.size).astype(np.float64), (8, 6))

    add("all_empty", np.zeros(4, np.int64), np.zeros(0, np.int64), np.zeros(0, np.float32), (3, 5))

    # rows summing (exactly) to zero or almost: a, -a pairs plus a tiny term
    a = rng.uniform(-1, 1, 6).astype(np.float32)
    data = np.concatenate([a, -a, [np.float32(1e-6)]]).astype(np.float32)
    cols = np.concatenate([np.arange(6), np.arange(6), [3]])
    add("cancel", [0, 12, 13], cols, data, (2, 6), x=np.ones(6, np.float32))

    lens = np.array([2, 0, 3])
    ip = np.concatenate([[0], np.cumsum(lens)])
    cols = np.array([0, 3, 1, 1, 2])
    cdata = (rng.uniform(-1, 1, 5) + 1j * rng.uniform(-1, 1, 5)).astype(np.complex128)
    cx = (rng.uniform(-1, 1, 4) + 1j * rng.uniform(-1, 1, 4)).astype(np.complex128)
    add("complex", ip, cols, cdata, (3, 4), x=cx)

    out = os.path.join(HERE, "csr_golden.npz")
    np.savez_compressed(out, **cases)
    print(out, len(cases), "arrays")


if __name__ == "__main__":
    main()
