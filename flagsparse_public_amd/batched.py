"""Many uneven fp32 CSR SpMV / SpMM problems in one persistent launch.

``BatchedCsr`` takes a list of matrices (each with an optional SpMV operand x
and an optional SpMM operand B), validates them, packs one device record per
matrix, and builds a device work list whose items carry about the same
number of bytes (entries weighted by the dense columns they gather for SpMM,
rows cut at most every ``max_rows``).  ``run()`` launches the persistent
kernel once; ``capture()`` records that launch into a HIP graph
(``torch.cuda.CUDAGraph``) so a replay costs one graph launch.  The work
queue resets itself at the end of every launch, so the graph needs no memset
node.

Rows use the single-matrix kernels' per-row procedures: with the same lane
group width a row's value is bit-identical to ``flagsparse_spmv_csr`` /
``flagsparse_spmm_csr_opt`` on that matrix (rows short enough for the
short-row role).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .rowsplit import rowgroup_spmv_schedule
from .sparse_operations.spmv_csr import _prepare_spmv_csr_matrix


FsBatchEntry = _lib.FsBatchEntry


def _lib_batch():
    return _lib.load()


SPMV, SPMM = 0, 1


def _cuts(indptr, weight_per_entry, target, max_rows, long_row=None):
    """Row cut points of one matrix: a cut at the row holding every multiple
    of `target` weighted entries, at least every `max_rows` rows, and around
    every row longer than `long_row` (which then forms an item of its own)."""
    n = indptr.numel() - 1
    dev = indptr.device
    nnz = int(indptr[-1].item())
    step = max(1, int(target // max(weight_per_entry, 1e-9)))
    marks = (torch.arange(step, nnz, step, device=dev, dtype=torch.int64) if nnz > step
             else torch.zeros(0, dtype=torch.int64, device=dev))
    by_nnz = torch.searchsorted(indptr[1:].to(torch.int64), marks, right=True)
    by_rows = (torch.arange(max_rows, n, max_rows, device=dev, dtype=torch.int64) if n > max_rows
               else torch.zeros(0, dtype=torch.int64, device=dev))
    parts = [torch.zeros(1, dtype=torch.int64, device=dev), by_nnz, by_rows,
             torch.full((1,), n, dtype=torch.int64, device=dev)]
    if long_row is not None and n > 0:
        lens = indptr[1:] - indptr[:-1]
        long = torch.nonzero(lens > long_row, as_tuple=False).flatten().to(torch.int64)
        parts += [long, long + 1]
    cuts = torch.cat(parts)
    cuts = torch.unique(cuts)  # sorted
    return cuts[(cuts >= 0) & (cuts <= n)]


class BatchedCsr:
    """See module docstring.  ``matrices`` is a list of dicts with keys
    data, indices, indptr, shape and optionally x (SpMV), B (SpMM) and
    group_lanes."""

    def __init__(self, matrices, item_bytes=96 * 1024, max_rows=2048, long_row=4096,
                 blocks_per_cu=8, validate=True):
        if not matrices:
            raise ValueError("empty batch")
        dev = matrices[0]["data"].device
        self.device = dev
        self.long_row = int(long_row)
        self.entries = (FsBatchEntry * len(matrices))()
        self.keep = []  # tensors referenced by device records
        self.y, self.C = [], []
        self.group_lanes = []  # SpMV lanes per row of each matrix
        items = []
        bytes_total = 0
        for i, m in enumerate(matrices):
            data, indices, indptr = m["data"], m["indices"], m["indptr"]
            shape = (int(m["shape"][0]), int(m["shape"][1]))
            if data.dtype != torch.float32:
                raise TypeError("batched path takes float32 matrices")
            if validate:
                data, indices, indptr, n_rows, n_cols, _, _ = _prepare_spmv_csr_matrix(
                    data, indices, indptr, shape)
            else:
                n_rows, n_cols = shape
            if indices.dtype != torch.int32:
                indices = indices.to(torch.int32)
            x, B = m.get("x"), m.get("B")
            y = torch.zeros(n_rows, dtype=torch.float32, device=dev) if x is not None else None
            if x is not None and (x.dtype != torch.float32 or x.numel() != n_cols or not x.is_contiguous()):
                raise ValueError(f"matrix {i}: x must be a contiguous float32 vector of n_cols")
            C = None
            n_dense = 0
            if B is not None:
                if B.dtype != torch.float32 or B.shape[0] != n_cols or B.ndim != 2:
                    raise ValueError(f"matrix {i}: B must be float32 of shape (n_cols, N)")
                B = B.contiguous()
                n_dense = int(B.shape[1])
                C = torch.zeros((n_rows, n_dense), dtype=torch.float32, device=dev)
            nnz = data.numel()
            s = m.get("group_lanes") or rowgroup_spmv_schedule(n_rows, nnz).group_lanes
            e = self.entries[i]
            e.A = _lib.make_csr(data, indices, indptr, n_rows, n_cols)
            e.x, e.y = _lib.ptr(x), _lib.ptr(y)
            e.B, e.C = _lib.ptr(B), _lib.ptr(C)
            e.n_dense, e.ldb, e.ldc = n_dense, n_dense, n_dense
            e.group_lanes = int(s)
            self.group_lanes.append(int(s))
            self.keep.extend([data, indices, indptr, x, B])
            self.y.append(y)
            self.C.append(C)
            ip64 = indptr.to(torch.int64)
            if x is not None and n_rows:
                cuts = _cuts(ip64, 12.0, item_bytes, max_rows, self.long_row)
                items.append(self._items(i, SPMV, cuts))
                bytes_total += 12 * nnz + 8 * n_rows
            if B is not None and n_rows and n_dense:
                cuts = _cuts(ip64, 8.0 + 4.0 * n_dense, item_bytes, max_rows)
                items.append(self._items(i, SPMM, cuts))
                bytes_total += (8 + 4 * n_dense) * nnz + 4 * n_dense * n_rows
        lib = _lib_batch()
        rec = int(lib.fs_batch_matrix_bytes())
        host = (ctypes.c_uint8 * (rec * len(matrices)))()
        _lib.check(lib.fs_batch_pack(self.entries, len(matrices), ctypes.addressof(host)),
                   "fs_batch_pack")
        self.packed = torch.frombuffer(bytearray(host), dtype=torch.uint8).to(dev)
        self.items = torch.cat(items) if items else torch.zeros((0, 3), dtype=torch.int64, device=dev)
        self.n_items = int(self.items.shape[0])
        self.n_matrices = len(matrices)
        self.queue = torch.zeros(2, dtype=torch.int32, device=dev)
        props = torch.cuda.get_device_properties(dev)
        self.grid = int(props.multi_processor_count) * int(blocks_per_cu)
        self.bytes = bytes_total
        self._graph = None

    @staticmethod
    def _items(matrix, kind, cuts):
        r0, r1 = cuts[:-1], cuts[1:]
        keep = r1 > r0
        r0, r1 = r0[keep], r1[keep]
        head = torch.full_like(r0, matrix | (kind << 32))
        return torch.stack([head, r0, r1], dim=1)

    def run(self, alpha=1.0, beta=0.0):
        lib = _lib_batch()
        with torch.cuda.device(self.device):
            _lib.check(lib.fs_batch_run(self.packed.data_ptr(), self.n_matrices,
                                        self.items.data_ptr(), self.n_items, self.long_row,
                                        self.queue.data_ptr(), self.grid, float(alpha),
                                        float(beta), _lib.stream_handle(self.device)),
                       "fs_batch_run")

    def capture(self):
        """Record one run() into a HIP graph; returns it (replay with .replay())."""
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            self.run()  # warm (queue self-resets)
        torch.cuda.current_stream(self.device).wait_stream(s)
        with torch.cuda.graph(g):
            self.run()
        self._graph = g
        return g


def item_stats(batch: BatchedCsr):
    """(n_items, rows per item, max rows) summary of a work list."""
    it = batch.items.cpu().numpy()
    rows = it[:, 2] - it[:, 1]
    return {"n_items": int(it.shape[0]), "mean_rows": float(rows.mean()) if rows.size else 0.0,
            "max_rows": int(rows.max()) if rows.size else 0}


__all__ = ["BatchedCsr", "item_stats", "SPMV", "SPMM", "np"]
