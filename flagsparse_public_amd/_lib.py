"""ctypes binding of libflagsparse_hip.so (the C ABI in include/flagsparse_hip.h).

This module is the only place that touches the shared library.  It maps
torch dtypes to the ABI's value/index codes, packs device pointers into the
``fs_csr`` / ``fs_plan`` structs, passes torch's current HIP stream, and turns
``FS_ERR_*`` statuses into the Python exceptions the reference raises
(ValueError / TypeError / IndexError / RuntimeError).

There is no fallback: if the library is missing or cannot be loaded, every
entry point raises ``RuntimeError``.
"""

from __future__ import annotations

import ctypes
import os
import threading

import torch

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG_DIR, "libflagsparse_hip.so")

FS_OK, FS_ERR_VALUE, FS_ERR_TYPE, FS_ERR_INDEX, FS_ERR_HIP, FS_ERR_UNSUPPORTED = range(6)
FS_F16, FS_BF16, FS_F32, FS_F64, FS_C32, FS_C64, FS_C128 = range(7)
FS_I32, FS_I64 = 0, 1
FS_ACC_NATIVE, FS_ACC_WIDE = 0, 1
FS_STORE_BF16X2, FS_STORE_BF16, FS_STORE_F16_ROWSCALE = 1, 2, 3
FS_STATS_LEN = 16
FS_TILE_ROWS, FS_TILE_ENTRIES = 0, 1
FS_XG_TILE_MAJOR, FS_XG_WINDOW_MAJOR = 0, 1
FS_NOALLOC_SCRATCH_BYTES = 4 << 20  # fs_spmv_csr_noalloc's per-GPU partials scratch

_COMPLEX32 = getattr(torch, "complex32", None) or getattr(torch, "chalf", None)

VALUE_CODE = {
    torch.float16: FS_F16,
    torch.bfloat16: FS_BF16,
    torch.float32: FS_F32,
    torch.float64: FS_F64,
    torch.complex64: FS_C64,
    torch.complex128: FS_C128,
}
if _COMPLEX32 is not None:
    VALUE_CODE[_COMPLEX32] = FS_C32
INDEX_CODE = {torch.int32: FS_I32, torch.int64: FS_I64}


class FsCsr(ctypes.Structure):
    _fields_ = [
        ("value_type", ctypes.c_int32),
        ("index_type", ctypes.c_int32),
        ("ptr_type", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("n_rows", ctypes.c_int64),
        ("n_cols", ctypes.c_int64),
        ("nnz", ctypes.c_int64),
        ("row_ptr", ctypes.c_void_p),
        ("col_idx", ctypes.c_void_p),
        ("values", ctypes.c_void_p),
    ]


class FsPlan(ctypes.Structure):
    _fields_ = [
        ("accumulate", ctypes.c_int32),
        ("group_lanes", ctypes.c_int32),
        ("xcd_remap", ctypes.c_int32),
        ("unroll", ctypes.c_int32),
        ("short_max", ctypes.c_int64),
        ("chunk_nnz", ctypes.c_int64),
        ("n_long", ctypes.c_int64),
        ("long_rows", ctypes.c_void_p),
        ("long_chunk_ptr", ctypes.c_void_p),
        ("n_chunks", ctypes.c_int64),
        ("chunk_long", ctypes.c_void_p),
        ("partials", ctypes.c_void_p),
        ("partials_bytes", ctypes.c_int64),
        ("tile_nnz", ctypes.c_int64),
        ("n_tiles", ctypes.c_int64),
        ("tile_start", ctypes.c_void_p),
        ("tile_mode", ctypes.c_int32),
        ("n_groups", ctypes.c_int32),
        ("group_blocks", ctypes.c_int64),
        ("group_tiles", ctypes.c_void_p),
        ("row_sums", ctypes.c_void_p),
        ("tile_entry", ctypes.c_void_p),
        ("x_perm", ctypes.c_void_p),
        ("xg_tile_run", ctypes.c_void_p),
        ("xg_batch_run", ctypes.c_void_p),
        ("xg_batch_mask", ctypes.c_void_p),
        ("xg_run_off", ctypes.c_void_p),
        ("xg_n_runs", ctypes.c_int64),
    ]


class FsBatchEntry(ctypes.Structure):
    _fields_ = [
        ("A", FsCsr),
        ("x", ctypes.c_void_p),
        ("y", ctypes.c_void_p),
        ("B", ctypes.c_void_p),
        ("C", ctypes.c_void_p),
        ("n_dense", ctypes.c_int64),
        ("ldb", ctypes.c_int64),
        ("ldc", ctypes.c_int64),
        ("group_lanes", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class FsSlices(ctypes.Structure):
    _fields_ = [
        ("value_type", ctypes.c_int32),
        ("n_slices", ctypes.c_int32),
        ("n_rows", ctypes.c_int64),
        ("n_compact", ctypes.c_int64),
        ("compact_row", ctypes.c_void_p),
        ("bounds", ctypes.c_void_p),
        ("rows_per_block", ctypes.c_int32),
        ("z_type", ctypes.c_int32),
        ("row_mask", ctypes.c_void_p),
        ("fp32_sums", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class FsXwin(ctypes.Structure):
    _fields_ = [
        ("value_type", ctypes.c_int32),
        ("window", ctypes.c_int32),
        ("n_cols", ctypes.c_int64),
        ("nnz", ctypes.c_int64),
        ("n_runs", ctypes.c_int64),
        ("n_parts", ctypes.c_int64),
        ("wcol", ctypes.c_void_p),
        ("batch_run", ctypes.c_void_p),
        ("batch_mask", ctypes.c_void_p),
        ("run_off", ctypes.c_void_p),
        ("part_window", ctypes.c_void_p),
        ("part_entry", ctypes.c_void_p),
        ("window_entry", ctypes.c_void_p),
        ("order", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class FsMtxInfo(ctypes.Structure):
    _fields_ = [
        ("n_rows", ctypes.c_int64),
        ("n_cols", ctypes.c_int64),
        ("nnz", ctypes.c_int64),
        ("field", ctypes.c_int32),
        ("symmetry", ctypes.c_int32),
    ]


class FsMlp(ctypes.Structure):
    _fields_ = [
        ("in_dim", ctypes.c_int32),
        ("hidden", ctypes.c_int32),
        ("out_dim", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("params", ctypes.c_void_p),
    ]


_lib = None
_lib_lock = threading.Lock()

# Every symbol the header declares, with its ctypes signature.
_SIGNATURES = {
    "fs_xwin_max_window": (ctypes.c_int64, [ctypes.c_int32]),
    "fs_xwin_gather": (ctypes.c_int, [ctypes.POINTER(FsXwin), ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p]),
    "fs_partials_bytes": (ctypes.c_int64, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64]),
    "fs_tile_carry_bytes": (ctypes.c_int64, [ctypes.c_int32, ctypes.c_int64]),
    "fs_sliced_combine": (ctypes.c_int, [ctypes.POINTER(FsSlices), ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_void_p]),
    "fs_spmv_csr": (ctypes.c_int, [ctypes.POINTER(FsCsr), ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_double, ctypes.c_double, ctypes.POINTER(FsPlan),
                                   ctypes.c_void_p]),
    "fs_spmv_csr_noalloc": (ctypes.c_int, [ctypes.POINTER(FsCsr), ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_double, ctypes.c_double, ctypes.POINTER(FsPlan),
                                           ctypes.c_void_p]),
    "fs_spmm_csr": (ctypes.c_int, [ctypes.POINTER(FsCsr), ctypes.c_void_p, ctypes.c_int64,
                                   ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                   ctypes.c_double, ctypes.c_double, ctypes.POINTER(FsPlan),
                                   ctypes.c_int32, ctypes.c_void_p]),
    "fs_csr_analyze": (ctypes.c_int, [ctypes.POINTER(FsCsr), ctypes.c_void_p, ctypes.c_void_p]),
    "fs_split_values": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(FsCsr), ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "fs_spmv_csr_split": (ctypes.c_int, [ctypes.POINTER(FsCsr), ctypes.c_int32, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                         ctypes.POINTER(FsPlan), ctypes.c_void_p]),
    "fs_batch_matrix_bytes": (ctypes.c_int64, []),
    "fs_batch_pack": (ctypes.c_int, [ctypes.POINTER(FsBatchEntry), ctypes.c_int64, ctypes.c_void_p]),
    "fs_batch_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                    ctypes.c_int64, ctypes.c_void_p, ctypes.c_int32, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_void_p]),
    "fs_mlp_param_count": (ctypes.c_int64, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "fs_csr_features": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                       ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]),
    "fs_mlp_forward": (ctypes.c_int, [ctypes.POINTER(FsMlp), ctypes.c_void_p, ctypes.c_int64,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p]),
    "fs_mlp_workspace_bytes": (ctypes.c_int64, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                ctypes.c_int64]),
    "fs_mlp_backward": (ctypes.c_int, [ctypes.POINTER(FsMlp)] + [ctypes.c_void_p] * 6
                        + [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                           ctypes.c_int64, ctypes.c_void_p]),
    "fs_adam_step": (ctypes.c_int, [ctypes.c_void_p] * 4 + [ctypes.c_int64, ctypes.c_double,
                                                            ctypes.c_double, ctypes.c_double,
                                                            ctypes.c_double, ctypes.c_int64,
                                                            ctypes.c_void_p]),
    "fs_mlp_select": (ctypes.c_int, [ctypes.POINTER(FsMlp)] + [ctypes.c_void_p] * 6),
    "fs_sddmm_csr": (ctypes.c_int, [ctypes.POINTER(FsCsr), ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                    ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_void_p]),
    "fs_mtx_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(FsMtxInfo),
                                   ctypes.POINTER(ctypes.c_void_p)]),
    "fs_mtx_fill": (ctypes.c_int, [ctypes.c_void_p] * 4),
    "fs_mtx_free": (None, [ctypes.c_void_p]),
    "fs_mtx_last_error": (ctypes.c_char_p, []),
    "fs_last_error": (ctypes.c_char_p, []),
    "fs_version": (ctypes.c_char_p, []),
}


def load():
    """Load the shared library (after torch, so torch's HIP runtime is reused)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m flagsparse_public_amd.build`")
        try:
            lib = ctypes.CDLL(LIB_PATH)
        except OSError as exc:
            raise RuntimeError(f"cannot load {LIB_PATH}: {exc}") from exc
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def exported_symbols():
    return tuple(_SIGNATURES)


def check(status: int, what: str = "") -> None:
    if status == FS_OK:
        return
    msg = load().fs_last_error().decode("utf-8", "replace")
    if what and not msg.startswith(what):
        msg = f"{what}: {msg}"
    if status == FS_ERR_VALUE:
        raise ValueError(msg)
    if status == FS_ERR_TYPE:
        raise TypeError(msg)
    if status == FS_ERR_INDEX:
        raise IndexError(msg)
    raise RuntimeError(msg)


def value_code(dtype) -> int:
    try:
        return VALUE_CODE[dtype]
    except KeyError:
        raise TypeError(f"unsupported value dtype {dtype}") from None


def index_code(dtype) -> int:
    try:
        return INDEX_CODE[dtype]
    except KeyError:
        raise TypeError(f"unsupported index dtype {dtype}") from None


def ptr(t) -> int | None:
    if t is None or t.numel() == 0:
        return None
    return t.data_ptr()


def make_csr(data, indices, indptr, n_rows, n_cols) -> FsCsr:
    return FsCsr(
        value_type=value_code(data.dtype),
        index_type=index_code(indices.dtype),
        ptr_type=index_code(indptr.dtype),
        reserved=0,
        n_rows=int(n_rows),
        n_cols=int(n_cols),
        nnz=int(data.numel()),
        row_ptr=ptr(indptr),
        col_idx=ptr(indices),
        values=ptr(data),
    )


class _Nothing:
    def __enter__(self):
        return None

    def __exit__(self, *exc):
        return False


_NOTHING = _Nothing()


def on_device(device):
    """Context making `device` current; a no-op when it already is (the common
    case: one GPU per process), which keeps the per-call host cost low."""
    idx = device.index if isinstance(device, torch.device) else int(device)
    if idx is None or idx == torch.cuda.current_device():
        return _NOTHING
    return torch.cuda.device(idx)


_RAW_STREAM = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def stream_handle(device) -> int:
    """Raw handle of `device`'s current HIP stream (what
    ``torch.cuda.current_stream(device).cuda_stream`` gives).  torch's raw
    query skips building a Stream object, about 2 us of the ~10 us host cost
    of a small prepared SpMV call; every caller already holds tensors on the
    device, so torch's GPU state is initialised."""
    if _RAW_STREAM is None:
        return torch.cuda.current_stream(device).cuda_stream
    if not isinstance(device, torch.device):
        device = torch.device(device) if isinstance(device, str) else torch.device("cuda", int(device))
    idx = device.index if device.index is not None else torch.cuda.current_device()
    return int(_RAW_STREAM(idx))


def analyze_device(csr: FsCsr, device):
    """Run fs_csr_analyze; returns the FS_STATS_LEN statistics on the device."""
    lib = load()
    stats = torch.empty(FS_STATS_LEN, dtype=torch.int64, device=device)
    with on_device(device):
        check(lib.fs_csr_analyze(ctypes.byref(csr), stats.data_ptr(), stream_handle(device)),
              "fs_csr_analyze")
    return stats


def analyze(csr: FsCsr, device) -> list[int]:
    """Run fs_csr_analyze and copy the FS_STATS_LEN statistics to the host."""
    return analyze_device(csr, device).cpu().tolist()


def features(stats, n_rows: int, n_cols: int, nnz: int, device):
    """16 selector features (device float32) from device statistics."""
    f = torch.empty(16, dtype=torch.float32, device=device)
    with on_device(device):
        check(load().fs_csr_features(stats.data_ptr(), int(n_rows), int(n_cols), int(nnz),
                                     f.data_ptr(), stream_handle(device)), "fs_csr_features")
    return f


def mlp_struct(params, in_dim: int, hidden: int, out_dim: int) -> FsMlp:
    return FsMlp(in_dim=int(in_dim), hidden=int(hidden), out_dim=int(out_dim), reserved=0,
                 params=params.data_ptr())


def mlp_forward(net: FsMlp, X, H1, H2, Y, device) -> None:
    with on_device(device):
        check(load().fs_mlp_forward(ctypes.byref(net), ptr(X), int(X.shape[0]), ptr(H1), ptr(H2),
                                    ptr(Y), stream_handle(device)), "fs_mlp_forward")


def mlp_backward(net: FsMlp, X, H1, H2, Y, T, M, grad, loss, workspace, device) -> None:
    with on_device(device):
        check(load().fs_mlp_backward(ctypes.byref(net), ptr(X), ptr(H1), ptr(H2), ptr(Y), ptr(T),
                                     ptr(M), int(X.shape[0]), ptr(grad), ptr(loss), ptr(workspace),
                                     int(workspace.numel()), stream_handle(device)),
              "fs_mlp_backward")


def mlp_workspace_bytes(in_dim, hidden, out_dim, batch) -> int:
    return int(load().fs_mlp_workspace_bytes(in_dim, hidden, out_dim, batch))


def mlp_param_count(in_dim, hidden, out_dim) -> int:
    return int(load().fs_mlp_param_count(in_dim, hidden, out_dim))


def adam_step(params, grad, m, v, lr, beta1, beta2, eps, step, device) -> None:
    with on_device(device):
        check(load().fs_adam_step(ptr(params), ptr(grad), ptr(m), ptr(v), int(params.numel()),
                                  float(lr), float(beta1), float(beta2), float(eps), int(step),
                                  stream_handle(device)), "fs_adam_step")


def mlp_select(net: FsMlp, feats, mu, sd, choice, scores, device) -> None:
    with on_device(device):
        check(load().fs_mlp_select(ctypes.byref(net), ptr(feats), ptr(mu), ptr(sd), ptr(choice),
                                   ptr(scores), stream_handle(device)), "fs_mlp_select")


def spmv(csr: FsCsr, x, y, alpha: float, beta: float, plan: FsPlan, device) -> None:
    lib = load()
    with on_device(device):
        check(lib.fs_spmv_csr(ctypes.byref(csr), ptr(x), ptr(y), float(alpha), float(beta),
                              ctypes.byref(plan), stream_handle(device)), "fs_spmv_csr")


def spmm(csr: FsCsr, B, ldb: int, C, ldc: int, n_dense: int, alpha: float, beta: float,
         plan: FsPlan, col_tile: int, device) -> None:
    lib = load()
    with on_device(device):
        check(lib.fs_spmm_csr(ctypes.byref(csr), ptr(B), int(ldb), ptr(C), int(ldc),
                              int(n_dense), float(alpha), float(beta), ctypes.byref(plan),
                              int(col_tile), stream_handle(device)), "fs_spmm_csr")


def split_values(fmt: int, csr: FsCsr, hi, lo, row_scale, device) -> None:
    lib = load()
    with on_device(device):
        check(lib.fs_split_values(int(fmt), ctypes.byref(csr), ptr(hi), ptr(lo), ptr(row_scale),
                                  stream_handle(device)), "fs_split_values")


def spmv_split(csr: FsCsr, fmt: int, hi, lo, row_scale, x, y, alpha: float, beta: float,
               plan: FsPlan, device) -> None:
    lib = load()
    with on_device(device):
        check(lib.fs_spmv_csr_split(ctypes.byref(csr), int(fmt), ptr(hi), ptr(lo), ptr(row_scale),
                                    ptr(x), ptr(y), float(alpha), float(beta), ctypes.byref(plan),
                                    stream_handle(device)), "fs_spmv_csr_split")


def partials_bytes(value_type: int, accumulate: int, n_chunks: int, n_dense: int = 1) -> int:
    return int(load().fs_partials_bytes(value_type, accumulate, n_chunks, n_dense))


def sliced_combine(sl: FsSlices, z, y, alpha: float, beta: float, device) -> None:
    check(load().fs_sliced_combine(ctypes.byref(sl), ptr(z), ptr(y), float(alpha), float(beta),
                                   stream_handle(device)), "fs_sliced_combine")


def xwin_max_window(value_type: int, device) -> int:
    with on_device(device):
        return int(load().fs_xwin_max_window(int(value_type)))


def xwin_gather(layout: FsXwin, x, xg, device) -> None:
    with on_device(device):
        check(load().fs_xwin_gather(ctypes.byref(layout), ptr(x), ptr(xg), stream_handle(device)),
              "fs_xwin_gather")


def sddmm(csr: FsCsr, row_ids, x, y, k_dim: int, out, out_type: int, accumulate: int,
          alpha: float, beta: float, device) -> None:
    lib = load()
    with on_device(device):
        check(lib.fs_sddmm_csr(ctypes.byref(csr), ptr(row_ids), ptr(x), ptr(y), int(k_dim),
                               ptr(out), int(out_type), int(accumulate), float(alpha),
                               float(beta), stream_handle(device)), "fs_sddmm_csr")


def tile_carry_bytes(value_type: int, n_tiles: int) -> int:
    return int(load().fs_tile_carry_bytes(value_type, n_tiles))


def version() -> str:
    return load().fs_version().decode()
