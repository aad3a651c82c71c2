"""MatrixMarket loader (csrc/mtx_io.cpp, mtx.py) against SciPy's reader and
hand-computed cases; the reading rules are those of the reference harness
(tests/test_spmv.py:103-176).  CPU only."""

import io

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp
import torch

from flagsparse_public_amd import mtx


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def _scipy_csr(path):
    a = scipy.io.mmread(path)
    a = sp.csr_matrix(a)
    a.sum_duplicates()
    a.sort_indices()
    return a


def _random_mtx(rng, n, m, k, field, symmetry, dup=True):
    """Entries in random order, with repeats; lower triangle for symmetric kinds."""
    rows = rng.integers(0, n, k)
    cols = rng.integers(0, m, k)
    if symmetry != "general":
        lo = np.minimum(rows, cols)
        hi = np.maximum(rows, cols)
        rows, cols = hi, lo
        if symmetry == "skew-symmetric":
            keep = rows != cols
            rows, cols = rows[keep], cols[keep]
    if dup and rows.size > 4:
        rows = np.concatenate([rows, rows[:3]])
        cols = np.concatenate([cols, cols[:3]])
    vals = rng.uniform(-1, 1, rows.size)
    out = io.StringIO()
    out.write(f"%%MatrixMarket matrix coordinate {field} {symmetry}\n% a comment\n%\n\n")
    out.write(f"{n} {m} {rows.size}\n")
    for r, c, v in zip(rows, cols, vals):
        if field == "pattern":
            out.write(f"{r + 1} {c + 1}\n")
        elif field == "integer":
            out.write(f"{r + 1} {c + 1} {int(round(v * 100))}\n")
        elif field == "complex":
            out.write(f"{r + 1} {c + 1} {float(v)!r} {float(-v)!r}\n")
        else:
            out.write(f"{r + 1} {c + 1} {float(v)!r}\n")
    return out.getvalue()


@pytest.mark.parametrize("field", ["real", "integer", "pattern", "complex"])
@pytest.mark.parametrize("symmetry", ["general", "symmetric", "skew-symmetric", "hermitian"])
def test_loader_matches_scipy(tmp_path, field, symmetry):
    if symmetry == "hermitian" and field != "complex":
        pytest.skip("hermitian files are complex")
    if symmetry == "skew-symmetric" and field == "pattern":
        pytest.skip("no skew-symmetric pattern files")
    rng = np.random.default_rng(hash((field, symmetry)) % 2**32)
    n = m = 60
    path = _write(tmp_path, "a.mtx", _random_mtx(rng, n, m, 400, field, symmetry))
    ip, ix, vals, shape, meta = mtx.read_mtx(path)
    assert shape == (n, m) and meta == {"field": field, "symmetry": symmetry}
    ref = _scipy_csr(path)
    if field == "complex":
        if symmetry == "hermitian":
            # the reference mirrors the real part unchanged (tests/test_spmv.py:160-165):
            # the real part of a hermitian matrix is symmetric, so this equals Re(A)
            ref = sp.csr_matrix(ref.real)
        else:
            ref = sp.csr_matrix(ref.real)
    if field == "pattern":
        # repeated pattern entries count once each in the reference (1.0 added per line)
        lines = [l.split() for l in open(path) if l.strip() and not l.startswith("%")][1:]
        r = np.array([int(a) - 1 for a, _ in lines])
        c = np.array([int(b) - 1 for _, b in lines])
        if symmetry == "symmetric":
            off = r != c
            r, c = np.concatenate([r, c[off]]), np.concatenate([c, r[off]])
        ref = sp.csr_matrix((np.ones(r.size), (r, c)), shape=(n, m))
        ref.sum_duplicates()
        ref.sort_indices()
    assert np.array_equal(ip, ref.indptr) and np.array_equal(ix, ref.indices)
    assert np.allclose(vals, ref.data, rtol=1e-15, atol=1e-15)


def test_loader_hand_cases(tmp_path):
    # out-of-range entries are dropped, only the declared number of lines is
    # read, repeats are summed in file order, columns come out sorted
    text = ("%%MatrixMarket matrix coordinate real general\n"
            "3 4 6\n"
            "1 3 1.5\n"
            "3 1 2.0\n"
            "1 1 -1.0\n"
            "4 1 9.0\n"      # row 4 > 3: dropped
            "1 3 0.25\n"     # repeat of (0, 2)
            "2 5 7.0\n"      # column 5 > 4: dropped
            "2 2 100.0\n")   # beyond the 6 declared entries: not read
    ip, ix, vals, shape, _ = mtx.read_mtx(_write(tmp_path, "h.mtx", text))
    assert shape == (3, 4)
    assert ip.tolist() == [0, 2, 2, 3]
    assert ix.tolist() == [0, 2, 0]
    assert vals.tolist() == [-1.0, 1.75, 2.0]
    # skew-symmetric: the mirror is negated, the diagonal is not mirrored
    text = ("%%MatrixMarket matrix coordinate real skew-symmetric\n"
            "3 3 2\n2 1 4.0\n3 3 1.0\n")
    ip, ix, vals, _, _ = mtx.read_mtx(_write(tmp_path, "s.mtx", text))
    assert ip.tolist() == [0, 1, 2, 3] and ix.tolist() == [1, 0, 2]
    assert vals.tolist() == [-4.0, 4.0, 1.0]
    # repeats are summed in file order (fp64): (1e16 + 1) - 1e16 = 0, not 1
    text = ("%%MatrixMarket matrix coordinate real general\n"
            "1 2 3\n1 2 1e16\n1 2 1.0\n1 2 -1e16\n")
    _, _, vals, _, _ = mtx.read_mtx(_write(tmp_path, "o.mtx", text))
    assert vals.tolist() == [0.0]
    # empty and header-only
    ip, ix, vals, shape, _ = mtx.read_mtx(_write(tmp_path, "e.mtx", "%%MatrixMarket matrix coordinate real general\n5 5 0\n"))
    assert shape == (5, 5) and ip.tolist() == [0] * 6 and ix.size == 0


def test_loader_errors(tmp_path):
    with pytest.raises(ValueError, match="cannot open"):
        mtx.read_mtx(str(tmp_path / "missing.mtx"))
    with pytest.raises(ValueError, match="header"):
        mtx.read_mtx(_write(tmp_path, "c.mtx", "%%MatrixMarket matrix coordinate real general\n% only comments\n"))
    with pytest.raises(ValueError, match="bad value"):
        mtx.read_mtx(_write(tmp_path, "v.mtx", "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1\n"))


def test_load_to_torch(tmp_path):
    text = "%%MatrixMarket matrix coordinate real symmetric\n3 3 2\n2 1 0.1\n3 3 2.5\n"
    data, indices, indptr, shape = mtx.load_mtx_to_csr_torch(_write(tmp_path, "t.mtx", text),
                                                             dtype=torch.float32, device="cpu")
    assert shape == (3, 3) and data.dtype == torch.float32
    assert indices.dtype == torch.int64 and indptr.dtype == torch.int64
    assert indptr.tolist() == [0, 1, 2, 3] and indices.tolist() == [1, 0, 2]
    assert torch.equal(data, torch.tensor([0.1, 0.1, 2.5], dtype=torch.float32))
