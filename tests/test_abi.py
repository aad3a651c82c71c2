"""The C ABI library (CPU only): it loads, exports every function the header
declares, and rejects malformed calls with the documented status codes
before anything is launched."""

import ctypes
import os
import re

import pytest

from flagsparse_public_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "flagsparse_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fs_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert names, "no declarations found"
    for name in names:
        assert hasattr(lib, name), f"{name} declared in the header but not exported"
    # and the Python binding covers every declaration
    assert set(names) <= set(_lib.exported_symbols())


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of fs_csr / fs_plan have the C compiler's offsets."""
    import subprocess
    classes = {"fs_csr": _lib.FsCsr, "fs_plan": _lib.FsPlan, "fs_batch_entry": _lib.FsBatchEntry,
               "fs_mlp": _lib.FsMlp, "fs_slices": _lib.FsSlices,
               "fs_mtx_info": _lib.FsMtxInfo, "fs_xwin": _lib.FsXwin}
    fields = {st: [f for f, _ in cls._fields_] for st, cls in classes.items()}
    fields["fs_batch_item"] = ["matrix", "kind", "row_begin", "row_end"]
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void){"]
    for st, fs in fields.items():
        lines.append(f'printf("{st} %zu\\n", sizeof({st}));')
        for f in fs:
            lines.append(f'printf("{st}.{f} %zu\\n", offsetof({st}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", str(src), "-o", str(exe)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                 check=True).stdout.splitlines())
    for st, cls in classes.items():
        assert int(out[st]) == ctypes.sizeof(cls)
        for f in fields[st]:
            assert int(out[f"{st}.{f}"]) == getattr(cls, f).offset, (st, f)
    # work items are built as int64 triples: (matrix | kind << 32, row_begin, row_end)
    assert [int(out[f"fs_batch_item.{f}"]) for f in fields["fs_batch_item"]] == [0, 4, 8, 16]
    assert int(out["fs_batch_item"]) == 24


def test_version_and_partials_bytes():
    assert "gfx950" in _lib.version()
    assert _lib.partials_bytes(_lib.FS_F32, _lib.FS_ACC_NATIVE, 10, 1) == 80
    assert _lib.partials_bytes(_lib.FS_C128, _lib.FS_ACC_NATIVE, 10, 4) == 640
    assert _lib.partials_bytes(99, 0, 1, 1) == -1


def _csr(**kw):
    base = dict(value_type=_lib.FS_F32, index_type=_lib.FS_I32, ptr_type=_lib.FS_I32, reserved=0,
                n_rows=4, n_cols=4, nnz=4, row_ptr=16, col_idx=32, values=48)
    base.update(kw)
    return _lib.FsCsr(**base)


def _plan(**kw):
    base = dict(accumulate=0, group_lanes=8, xcd_remap=0, unroll=4, short_max=64, chunk_nnz=512,
                n_long=0, long_rows=None, long_chunk_ptr=None, n_chunks=0, chunk_long=None,
                partials=None, partials_bytes=0)
    base.update(kw)
    return _lib.FsPlan(**base)


def _spmv(csr, plan):
    lib = _lib.load()
    return lib.fs_spmv_csr(ctypes.byref(csr) if csr is not None else None, 64, 80, 1.0, 0.0,
                           ctypes.byref(plan) if plan is not None else None, None)


@pytest.mark.parametrize("csr,plan,code,msg", [
    (None, _plan(), _lib.FS_ERR_VALUE, "descriptor is null"),
    (_csr(value_type=42), _plan(), _lib.FS_ERR_TYPE, "value_type"),
    (_csr(index_type=7), _plan(), _lib.FS_ERR_TYPE, "FS_I32 or FS_I64"),
    (_csr(n_rows=-1), _plan(), _lib.FS_ERR_VALUE, "negative"),
    (_csr(row_ptr=None), _plan(), _lib.FS_ERR_VALUE, "row_ptr is null"),
    (_csr(n_cols=2**31), _plan(), _lib.FS_ERR_VALUE, "int32 column indices"),
    (_csr(), None, _lib.FS_ERR_VALUE, "plan is null"),
    (_csr(), _plan(accumulate=5), _lib.FS_ERR_VALUE, "accumulate"),
    (_csr(), _plan(short_max=-1), _lib.FS_ERR_VALUE, "short_max"),
    (_csr(), _plan(n_long=1), _lib.FS_ERR_VALUE, "metadata pointers"),
    (_csr(), _plan(n_long=1, long_rows=8, long_chunk_ptr=8, chunk_long=8,
                   n_chunks=1, chunk_nnz=100), _lib.FS_ERR_VALUE, "multiple of 64"),
    (_csr(), _plan(n_long=1, long_rows=8, long_chunk_ptr=8, chunk_long=8,
                   n_chunks=1, chunk_nnz=64, partials=8, partials_bytes=4),
     _lib.FS_ERR_VALUE, "partials buffer too small"),
    (_csr(), _plan(n_chunks=3), _lib.FS_ERR_VALUE, "chunks listed"),
])
def test_spmv_rejects_bad_arguments(csr, plan, code, msg):
    rc = _spmv(csr, plan)
    assert rc == code
    assert msg in _lib.load().fs_last_error().decode()


@pytest.mark.parametrize("tile", [
    dict(),                                    # row groups
    dict(tile_nnz=2048, n_tiles=1, tile_start=8, tile_mode=_lib.FS_TILE_ROWS),  # row tiles
    dict(n_long=1, long_rows=8, long_chunk_ptr=8, chunk_long=8, n_chunks=1, chunk_nnz=64,
         partials=8, partials_bytes=64),       # chunked long rows
])
def test_row_sums_only_with_entry_tiles(tile):
    """A plan that asks for raw row sums outside the entry-tile role is
    refused before launch, also when y is null (that role would store
    through y)."""
    lib = _lib.load()
    csr = _csr()
    plan = _plan(row_sums=128, **tile)
    for fn in (lambda: lib.fs_spmv_csr(ctypes.byref(csr), 64, None, 1.0, 0.0,
                                       ctypes.byref(plan), None),
               lambda: lib.fs_spmv_csr_split(ctypes.byref(csr), _lib.FS_STORE_BF16X2, 64, 72,
                                             None, 80, None, 1.0, 0.0, ctypes.byref(plan),
                                             None)):
        assert fn() == _lib.FS_ERR_VALUE
        assert "row_sums" in lib.fs_last_error().decode()


def test_batch_queue_bound():
    """The batched work queue counts in int32: item counts that would wrap it
    are refused before launch."""
    lib = _lib.load()
    rc = lib.fs_batch_run(8, 1, 16, 2**31 - 10, 4096, 24, 256, 1.0, 0.0, None)
    assert rc == _lib.FS_ERR_UNSUPPORTED
    assert "2^31" in lib.fs_last_error().decode()


def test_noalloc_scratch_bound():
    """fs_spmv_csr_noalloc: the scratch size the header states is the one the
    binding uses, and entry tiles whose carries exceed it are refused before
    anything is launched (the pointers are never dereferenced)."""
    text = open(HEADER).read()
    m = re.search(r"#define FS_NOALLOC_SCRATCH_BYTES \((\d+)u << (\d+)\)", text)
    assert m and int(m.group(1)) << int(m.group(2)) == _lib.FS_NOALLOC_SCRATCH_BYTES
    lib = _lib.load()
    n_tiles = _lib.FS_NOALLOC_SCRATCH_BYTES // 16 + 1  # fp32 carries: 16 bytes per tile
    csr = _csr(ptr_type=_lib.FS_I64, n_rows=n_tiles, n_cols=8, nnz=n_tiles * 1024)
    plan = _plan(short_max=0, chunk_nnz=64, tile_nnz=1024, n_tiles=n_tiles, tile_start=96,
                 tile_mode=_lib.FS_TILE_ENTRIES, tile_entry=112)
    rc = lib.fs_spmv_csr_noalloc(ctypes.byref(csr), 64, 80, 1.0, 0.0, ctypes.byref(plan), None)
    assert rc == _lib.FS_ERR_UNSUPPORTED
    assert "scratch" in lib.fs_last_error().decode()
    # the same plan one tile smaller fits (checked by the size query alone)
    assert _lib.tile_carry_bytes(_lib.FS_F32, n_tiles - 1) == _lib.FS_NOALLOC_SCRATCH_BYTES
    assert lib.fs_spmv_csr_noalloc(None, 64, 80, 1.0, 0.0, ctypes.byref(plan), None) == \
        _lib.FS_ERR_VALUE


def test_group_lanes_checked_before_launch():
    # a bad group width is reported without touching the device (n_rows > 0,
    # pointers are never dereferenced on the host)
    rc = _spmv(_csr(), _plan(group_lanes=3))
    assert rc == _lib.FS_ERR_VALUE
    assert "group_lanes" in _lib.load().fs_last_error().decode()


def test_status_to_exception_mapping():
    lib = _lib.load()
    lib.fs_spmv_csr(None, None, None, 1.0, 0.0, None, None)  # sets the message
    with pytest.raises(ValueError):
        _lib.check(_lib.FS_ERR_VALUE)
    with pytest.raises(TypeError):
        _lib.check(_lib.FS_ERR_TYPE)
    with pytest.raises(IndexError):
        _lib.check(_lib.FS_ERR_INDEX)
    with pytest.raises(RuntimeError):
        _lib.check(_lib.FS_ERR_HIP)
    _lib.check(_lib.FS_OK)


def test_spmm_and_analyze_reject_bad_arguments():
    lib = _lib.load()
    csr = _csr()
    rc = lib.fs_spmm_csr(ctypes.byref(csr), 64, 4, 80, 4, -1, 1.0, 0.0, ctypes.byref(_plan()), 0, None)
    assert rc == _lib.FS_ERR_VALUE and "n_dense" in lib.fs_last_error().decode()
    rc = lib.fs_spmm_csr(ctypes.byref(csr), 64, 4, 80, 2, 4, 1.0, 0.0, ctypes.byref(_plan()), 0, None)
    assert rc == _lib.FS_ERR_VALUE and "leading dimensions" in lib.fs_last_error().decode()
    rc = lib.fs_csr_analyze(ctypes.byref(csr), None, None)
    assert rc == _lib.FS_ERR_VALUE and "stats is null" in lib.fs_last_error().decode()
    bad = _csr(ptr_type=3)
    assert lib.fs_csr_analyze(ctypes.byref(bad), 8, None) == _lib.FS_ERR_TYPE


def test_dtype_codes_cover_reference_tables():
    import torch
    from flagsparse_public_amd.sparse_operations._common import SUPPORTED_VALUE_DTYPES
    from flagsparse_public_amd.sparse_operations.spmv_csr import SUPPORTED_SPMV_VALUE_DTYPES
    for d in SUPPORTED_VALUE_DTYPES + SUPPORTED_SPMV_VALUE_DTYPES:
        _lib.value_code(d)
    with pytest.raises(TypeError):
        _lib.value_code(torch.int8)


def _entry_plan(**kw):
    base = dict(short_max=0, chunk_nnz=64, tile_nnz=2048, n_tiles=1, tile_start=96,
                tile_mode=_lib.FS_TILE_ENTRIES, tile_entry=112, partials=8, partials_bytes=1024)
    base.update(kw)
    return _plan(**base)


@pytest.mark.parametrize("plan,msg", [
    # a gathered-x run table without x_perm, outside and inside entry tiles
    (_plan(xg_run_off=8), "run table is offered with x_perm only"),
    (_entry_plan(xg_tile_run=8, xg_batch_run=8, xg_batch_mask=8, xg_run_off=8, xg_n_runs=4),
     "run table is offered with x_perm only"),
    # window-major xg with a missing piece of its run table
    (_entry_plan(x_perm=8, xg_run_off=8, xg_batch_run=8, xg_batch_mask=8, xg_n_runs=4),
     "window-major xg needs"),
    (_entry_plan(x_perm=8, xg_run_off=8, xg_tile_run=8, xg_batch_run=8, xg_batch_mask=8,
                 xg_n_runs=0), "window-major xg needs"),
])
def test_spmv_rejects_bad_xg_run_tables(plan, msg):
    """Gathered-x plans whose window-major run table is incomplete are refused
    before launch (the pointers are never dereferenced)."""
    rc = _spmv(_csr(), plan)
    assert rc == _lib.FS_ERR_VALUE
    assert msg in _lib.load().fs_last_error().decode()


def _xwin(**kw):
    base = dict(value_type=_lib.FS_F32, window=1024, n_cols=4096, nnz=100, n_runs=10, n_parts=2,
                wcol=16, batch_run=32, batch_mask=48, run_off=64, part_window=80,
                part_entry=96, window_entry=112, order=_lib.FS_XG_WINDOW_MAJOR, reserved=0)
    base.update(kw)
    return _lib.FsXwin(**base)


@pytest.mark.parametrize("layout,xg,msg", [
    (_xwin(order=7), 256, "order must be"),
    (_xwin(value_type=_lib.FS_C64), 256, "fp32 / fp64"),
    (_xwin(window=1000 + 2), 256, "multiple of 4"),
    (_xwin(nnz=2**31), 256, "2^31 - 1"),
    (_xwin(wcol=18), 256, "16-byte aligned"),       # window-major: 16-byte wcol and xg
    (_xwin(), 260, "16-byte aligned"),
    (_xwin(order=_lib.FS_XG_TILE_MAJOR, run_off=None), 256, "null array"),
    (_xwin(order=_lib.FS_XG_TILE_MAJOR, n_runs=0), 256, "without runs"),
])
def test_xwin_gather_rejects_bad_layouts(layout, xg, msg):
    """fs_xwin_gather's argument checks, all before any launch; a
    window-major layout needs no run arrays (they may be null)."""
    lib = _lib.load()
    rc = lib.fs_xwin_gather(ctypes.byref(layout), 512, xg, None)
    assert rc in (_lib.FS_ERR_VALUE, _lib.FS_ERR_TYPE, _lib.FS_ERR_UNSUPPORTED)
    assert msg in lib.fs_last_error().decode()
