// status.hip -- error reporting, plan validation and small queries of the C ABI.
#include <string.h>

#include "fs_host.hpp"

namespace fs {

static thread_local char g_last_error[512] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

int64_t partial_elem_bytes(int32_t value_type, int32_t accumulate) {
  (void)accumulate;  // every chunk accumulator is 8 bytes except complex128
  switch (value_type) {
    case FS_C128: return 16;
    case FS_F16: case FS_BF16: case FS_F32: case FS_F64: case FS_C32: case FS_C64: return 8;
    default: return -1;
  }
}

int check_csr(const fs_csr* A, const char* who) {
  if (!A) return fail(FS_ERR_VALUE, "%s: matrix descriptor is null", who);
  if (A->value_type < FS_F16 || A->value_type > FS_C128)
    return fail(FS_ERR_TYPE, "%s: unsupported value_type %d", who, A->value_type);
  if ((A->index_type != FS_I32 && A->index_type != FS_I64) ||
      (A->ptr_type != FS_I32 && A->ptr_type != FS_I64))
    return fail(FS_ERR_TYPE, "%s: index and row-pointer types must be FS_I32 or FS_I64", who);
  if (A->n_rows < 0 || A->n_cols < 0 || A->nnz < 0)
    return fail(FS_ERR_VALUE, "%s: negative shape or nnz", who);
  if (A->n_rows > 0 && !A->row_ptr) return fail(FS_ERR_VALUE, "%s: row_ptr is null", who);
  if (A->nnz > 0 && (!A->col_idx || !A->values))
    return fail(FS_ERR_VALUE, "%s: col_idx or values is null", who);
  if (A->index_type == FS_I32 && A->n_cols > 0x7fffffffLL)
    return fail(FS_ERR_VALUE, "%s: int32 column indices cannot address %lld columns", who,
                (long long)A->n_cols);
  if (A->ptr_type == FS_I32 && A->nnz > 0x7fffffffLL)
    return fail(FS_ERR_VALUE, "%s: int32 row pointers cannot address %lld entries", who,
                (long long)A->nnz);
  return FS_OK;
}

int check_plan(const fs_plan* P, const fs_csr* A, int64_t n_dense, const char* who) {
  if (!P) return fail(FS_ERR_VALUE, "%s: plan is null", who);
  if (P->accumulate != FS_ACC_NATIVE && P->accumulate != FS_ACC_WIDE)
    return fail(FS_ERR_VALUE, "%s: accumulate must be FS_ACC_NATIVE or FS_ACC_WIDE", who);
  if (P->short_max < 0) return fail(FS_ERR_VALUE, "%s: short_max must be >= 0", who);
  if (P->n_long < 0 || P->n_chunks < 0) return fail(FS_ERR_VALUE, "%s: negative long-row counts", who);
  if (P->n_long > 0) {
    if (!P->long_rows || !P->long_chunk_ptr || !P->chunk_long)
      return fail(FS_ERR_VALUE, "%s: long-row metadata pointers are null", who);
    if (P->chunk_nnz <= 0 || P->chunk_nnz % kWave != 0)
      return fail(FS_ERR_VALUE, "%s: chunk_nnz must be a positive multiple of 64", who);
    if (P->n_chunks < P->n_long)
      return fail(FS_ERR_VALUE, "%s: fewer chunks than long rows", who);
    const int64_t need = partial_elem_bytes(A->value_type, P->accumulate) * P->n_chunks *
                         (n_dense > 0 ? n_dense : 1);
    if (!P->partials || P->partials_bytes < need)
      return fail(FS_ERR_VALUE, "%s: partials buffer too small (%lld < %lld bytes)", who,
                  (long long)P->partials_bytes, (long long)need);
  } else if (P->n_chunks != 0) {
    return fail(FS_ERR_VALUE, "%s: chunks listed without long rows", who);
  }
  if (P->tile_nnz < 0 || P->n_tiles < 0) return fail(FS_ERR_VALUE, "%s: negative tile fields", who);
  if (P->tile_mode != FS_TILE_ROWS && P->tile_mode != FS_TILE_ENTRIES)
    return fail(FS_ERR_VALUE, "%s: tile_mode must be FS_TILE_ROWS or FS_TILE_ENTRIES", who);
  // only the entry-tile role stores raw row sums or reads gathered x
  if (P->x_perm && !(P->tile_nnz > 0 && P->tile_mode == FS_TILE_ENTRIES))
    return fail(FS_ERR_VALUE, "%s: x_perm is offered for entry tiles only", who);
  if ((P->xg_run_off || P->xg_tile_run || P->xg_batch_run || P->xg_batch_mask) && !P->x_perm)
    return fail(FS_ERR_VALUE, "%s: an xg run table is offered with x_perm only", who);
  if (P->row_sums && !(P->tile_nnz > 0 && P->tile_mode == FS_TILE_ENTRIES))
    return fail(FS_ERR_VALUE, "%s: row_sums are offered for entry tiles only", who);
  if (P->tile_nnz > 0 && P->tile_mode == FS_TILE_ENTRIES) {
    if (n_dense != 1) return fail(FS_ERR_VALUE, "%s: entry tiles are for SpMV", who);
    if (P->n_long != 0) return fail(FS_ERR_VALUE, "%s: entry tiles take no long-row lists", who);
    const int64_t least = A->nnz > 0 ? (A->nnz + P->tile_nnz - 1) / P->tile_nnz : 1;
    if (P->n_tiles < least)
      return fail(FS_ERR_VALUE, "%s: entry tiles need n_tiles >= %lld, got %lld", who,
                  (long long)least, (long long)P->n_tiles);
    if (!P->tile_start || !P->tile_entry)
      return fail(FS_ERR_VALUE, "%s: entry tiles need tile_start and tile_entry", who);
    if (P->n_groups < 0 || (P->n_groups > 0 && (!P->group_tiles || P->group_blocks < 0)))
      return fail(FS_ERR_VALUE, "%s: bad block groups", who);
    if (A->value_type != FS_F32 && A->value_type != FS_F64 && A->value_type != FS_C64 &&
        A->value_type != FS_C128 && P->row_sums)
      return fail(FS_ERR_VALUE, "%s: row sums are offered for fp32 / fp64 / complex values", who);
    if (P->x_perm && ((A->value_type != FS_F32 && A->value_type != FS_F64) || P->row_sums))
      return fail(FS_ERR_VALUE, "%s: gathered x (x_perm) is offered for fp32 / fp64 values "
                  "without row_sums", who);
    if (P->xg_run_off && (!P->x_perm || !P->xg_tile_run || !P->xg_batch_run ||
                          !P->xg_batch_mask || P->xg_n_runs <= 0 || A->nnz <= 0))
      return fail(FS_ERR_VALUE, "%s: a window-major xg needs x_perm, xg_tile_run, "
                  "xg_batch_run, xg_batch_mask and xg_n_runs > 0", who);
    const int64_t need = fs_tile_carry_bytes(A->value_type, P->n_tiles);
    if (!P->partials || P->partials_bytes < need)
      return fail(FS_ERR_VALUE, "%s: carry buffer too small (%lld < %lld bytes)", who,
                  (long long)P->partials_bytes, (long long)need);
    return FS_OK;
  }
  if (P->tile_nnz > 0) {
    if (n_dense != 1) return fail(FS_ERR_VALUE, "%s: the tile role is for SpMV", who);
    if (P->short_max > P->tile_nnz)
      return fail(FS_ERR_VALUE, "%s: tile role needs short_max <= tile_nnz", who);
    if (A->n_rows > 0 && (P->n_tiles == 0 || !P->tile_start))
      return fail(FS_ERR_VALUE, "%s: tile role needs tile_start", who);
  }
  return FS_OK;
}

}  // namespace fs

extern "C" const char* fs_last_error(void) { return fs::g_last_error; }

extern "C" const char* fs_version(void) { return "flagsparse-hip 0.1 gfx950"; }

extern "C" int64_t fs_partials_bytes(int32_t value_type, int32_t accumulate, int64_t n_chunks,
                                     int64_t n_dense) {
  const int64_t e = fs::partial_elem_bytes(value_type, accumulate);
  if (e < 0 || n_chunks < 0) return -1;
  return e * n_chunks * (n_dense > 0 ? n_dense : 1);
}

extern "C" int64_t fs_tile_carry_bytes(int32_t value_type, int64_t n_tiles) {
  if (n_tiles < 0 || value_type < FS_F16 || value_type > FS_C128) return -1;
  const bool cplx = value_type == FS_C32 || value_type == FS_C64 || value_type == FS_C128;
  return 2 * n_tiles * (cplx ? 16 : 8);
}
