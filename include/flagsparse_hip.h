/*
 * flagsparse_hip.h -- C ABI of the gfx950 CSR SpMV / SpMM library
 * (libflagsparse_hip.so).
 *
 * Plain pointers and sizes only: every array argument is a device pointer
 * (HIP device memory), every call is enqueued on the caller's HIP stream
 * (`stream` is a hipStream_t passed as void*; NULL means the legacy default
 * stream), and calls are asynchronous on the device.  No entry point
 * allocates device memory except where stated: the caller owns inputs,
 * outputs, plans and workspaces.
 *
 * Every entry point returns FS_OK (0) or an FS_ERR_* status; fs_last_error()
 * returns the message of the calling thread's last failure.
 *
 * Each entry point replaces one kernel-launch routine of the reference
 * (paths relative to the reference repository, src/flagsparse/sparse_operations/):
 *   fs_spmv_csr        spmv_csr.py:664-708   _triton_spmv_csr_impl_prepared
 *                      spmv_csr.py:393-440   _triton_spmv_csr_impl_opt_prepared
 *   fs_spmm_csr        spmm_csr.py:525-599   _triton_spmm_csr_impl
 *                      spmm_csr.py:2034-2056 _triton_spmm_csr_impl_opt_prepared
 *                      spmm_csr_opt_alg2.py:771-798 _triton_spmm_csr_impl_opt_alg2_prepared
 *   fs_csr_analyze     spmv_csr.py:497-551 / spmm_csr.py:197-255, the
 *                      host-synchronising validation reductions
 *                      (indptr[0], indptr[-1], monotonicity, index range,
 *                      max row length), fused into one device pass.
 */
#ifndef FLAGSPARSE_HIP_H
#define FLAGSPARSE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------- status codes */
enum {
  FS_OK = 0,
  FS_ERR_VALUE = 1,       /* bad argument value or shape   -> ValueError  */
  FS_ERR_TYPE = 2,        /* unsupported dtype combination -> TypeError   */
  FS_ERR_INDEX = 3,       /* index out of range            -> IndexError  */
  FS_ERR_HIP = 4,         /* HIP runtime failure           -> RuntimeError */
  FS_ERR_UNSUPPORTED = 5  /* valid request this build cannot run          */
};

/* ------------------------------------------------------------ value types */
enum {
  FS_F16 = 0,
  FS_BF16 = 1,
  FS_F32 = 2,
  FS_F64 = 3,
  FS_C32 = 4,   /* complex32: (fp16 re, fp16 im) */
  FS_C64 = 5,   /* complex64: (fp32 re, fp32 im) */
  FS_C128 = 6   /* complex128 */
};

/* ------------------------------------------------------------ index types */
enum { FS_I32 = 0, FS_I64 = 1 };

/* ----------------------------------------------------- accumulation modes */
enum {
  FS_ACC_NATIVE = 0,  /* accumulate in the compute type of the values
                         (fp16/bf16 -> fp32); chunked long rows of fp32 use
                         compensated (TwoSum) fp32 summation                  */
  FS_ACC_WIDE = 1     /* fp32 values accumulated in fp64 and rounded once
                         (the reference's default path, spmv_csr.py:116-122) */
};

/* A CSR matrix resident in device memory. */
typedef struct fs_csr {
  int32_t value_type;  /* FS_F16 .. FS_C128                       */
  int32_t index_type;  /* FS_I32 / FS_I64: column indices         */
  int32_t ptr_type;    /* FS_I32 / FS_I64: row pointer            */
  int32_t reserved;
  int64_t n_rows;
  int64_t n_cols;
  int64_t nnz;
  const void *row_ptr; /* n_rows + 1 entries                      */
  const void *col_idx; /* nnz entries                             */
  const void *values;  /* nnz entries of value_type               */
} fs_csr;

/*
 * Row-split schedule shared by SpMV and SpMM.  Rows whose length is at most
 * short_max are processed in natural order by groups of group_lanes lanes
 * (SpMV) or by one lane group per row and column tile (SpMM).  Longer rows
 * are listed in long_rows and cut into chunks of chunk_nnz entries; one
 * wavefront reduces a chunk.  A row that fits one chunk is written by that
 * wavefront; otherwise every chunk stores its partial (one accumulator for
 * SpMV, one accumulator row of n_dense for SpMM) and a second small launch
 * sums each row's partials in chunk order (deterministic; no synchronisation
 * between workgroups inside a launch).  Size partials with fs_partials_bytes.
 */
/*
 * Tile role (SpMV, tile_nnz > 0): instead of one lane group per short row,
 * a workgroup takes a tile of consecutive rows [tile_start[t],
 * tile_start[t+1]), stages their row pointers and the products of all their
 * entries in LDS (entries loaded fully coalesced, several x gathers in flight
 * per lane) and sums each row in entry order.  Tiles hold at most 512 rows,
 * at most 2 * tile_nnz entries and either only rows of length <= short_max
 * or exactly one longer row (left to the chunk role); short_max <= tile_nnz.
 */
typedef struct fs_plan {
  int32_t accumulate;      /* FS_ACC_*                                     */
  int32_t group_lanes;     /* SpMV lanes per short row: 1,2,4,8,16,32,64   */
  int32_t xcd_remap;       /* 1: give each XCD a contiguous range of rows  */
  int32_t unroll;          /* entries in flight per lane: 1, 2, 4 or 8     */
  int64_t short_max;       /* longest row handled by the short-row role    */
  int64_t chunk_nnz;       /* entries per long-row chunk (multiple of 64)  */
  int64_t n_long;          /* number of listed long rows                   */
  const int64_t *long_rows;      /* [n_long] row ids, ascending            */
  const int64_t *long_chunk_ptr; /* [n_long + 1] first chunk of each row   */
  int64_t n_chunks;
  const int32_t *chunk_long;     /* [n_chunks] long-row ordinal of a chunk */
  void *partials;                /* chunk partial sums                     */
  int64_t partials_bytes;
  int64_t tile_nnz;              /* > 0: tile role for short rows (SpMV)   */
  int64_t n_tiles;
  const int64_t *tile_start;     /* [n_tiles + 1] first row of each tile   */
  int32_t tile_mode;             /* FS_TILE_ROWS or FS_TILE_ENTRIES        */
  int32_t n_groups;              /* FS_TILE_ENTRIES: 0, or block groups    */
  int64_t group_blocks;          /* blocks per group (largest tile range)  */
  const int64_t *group_tiles;    /* [n_groups + 1] tile range of a group   */
  void *row_sums;                /* FS_TILE_ENTRIES: if set, raw row sums  */
  const int64_t *tile_entry;     /* FS_TILE_ENTRIES: [n_tiles + 1]         */
  const uint16_t *x_perm;        /* FS_TILE_ENTRIES: if set, x is gathered */
  /* x_perm over a window-major xg (fs_xwin.order = FS_XG_WINDOW_MAJOR): the
     run table that finds each tile-order position's value in xg; all null
     (and xg_n_runs 0) for a tile-major xg.  Tile-order position d (the
     entries sorted by tile, then window, then CSR order) lies in run r =
     xg_batch_run[d / 64] + (set bits of xg_batch_mask[d / 64] at 1 .. d % 64)
     and takes xg[d + xg_run_off[r]]; tile t's positions lie in its runs
     xg_tile_run[t] .. xg_tile_run[t + 1] - 1 (at most tile_nnz of them). */
  const int32_t *xg_tile_run;    /* [n_tiles + 1]                          */
  const int32_t *xg_batch_run;   /* [ceil(nnz / 64)]                       */
  const uint64_t *xg_batch_mask; /* [ceil(nnz / 64)] bit i (i >= 1) set when
                                    position 64 b + i starts a run         */
  const int32_t *xg_run_off;     /* [xg_n_runs]                            */
  int64_t xg_n_runs;
} fs_plan;

/*
 * Tile modes of the SpMV (tile_nnz > 0):
 *   FS_TILE_ROWS     tiles of consecutive short rows holding about tile_nnz
 *                    entries; rows over short_max go to the listed long-row
 *                    chunks.
 *   FS_TILE_ENTRIES  tile t holds the entries [tile_entry[t],
 *                    tile_entry[t+1]) -- at most tile_nnz -- and owns the rows
 *                    tile_start[t] .. tile_start[t+1]-1, which start inside
 *                    that range (the last tile also owns rows starting at
 *                    nnz); tile_entry[0] = 0, tile_entry[n_tiles] = nnz,
 *                    tile_start[n_tiles] = n_rows, both nondecreasing.  The
 *                    planner cuts every tile_nnz entries and every 768 owned
 *                    rows.  No long-row lists (n_long = 0); partials = two
 *                    carries per tile (fs_tile_carry_bytes).  tile_nnz is
 *                    1024, 2048 or 4096 for fp32 / fp64 values with int32
 *                    columns, 2048 otherwise.
 *                    With n_groups = G > 0, the grid is G * group_blocks
 *                    blocks; block b serves group g = b % G and takes tile
 *                    group_tiles[g] + b / G (blocks past group_tiles[g + 1]
 *                    exit; the ranges must cover every tile once and
 *                    group_blocks must be the largest range): blocks are dispatched to the
 *                    8 XCDs round-robin, so G = 8 keeps each XCD on its own
 *                    contiguous tile range -- and its L2 on its own x columns
 *                    when the matrix is stored column-slice-major.
 *                    With row_sums set, y / alpha / beta are not used: row r's
 *                    sum is stored to row_sums[r] in the sum type (fp64; two
 *                    fp64 for complex), times its row scale for
 *                    FS_STORE_F16_ROWSCALE, for fs_sliced_combine to add up.
 *                    With x_perm set (fp32 / fp64 values, no row_sums), x is
 *                    not the dense vector but the nnz x values already
 *                    gathered per entry by fs_xwin_gather: tile t's values
 *                    lie at x[tile_entry[t] .. tile_entry[t+1]), in the
 *                    order x_perm gives -- entry j of tile t takes
 *                    x[tile_entry[t] + x_perm[j]] -- and col_idx is not read
 *                    (tile_nnz 1024, 2048 or 4096).  With xg_run_off also
 *                    set, x is window-major: tile position d = tile_entry[t]
 *                    + x_perm[j] is found at x[d + xg_run_off[run of d]].
 */
enum { FS_TILE_ROWS = 0, FS_TILE_ENTRIES = 1 };

/* Bytes of `partials` for FS_TILE_ENTRIES with n_tiles tiles. */
int64_t fs_tile_carry_bytes(int32_t value_type, int64_t n_tiles);

/* --------------------------------------------------------------- SDDMM */
/*
 * out[p] = alpha * sum_k x[row(p), k] * y[col(p), k] + beta * data[p] for every
 * entry p of the CSR pattern P (P->values = data, or null for beta = 0).
 * Replaces the reference's Triton SDDMM (sddmm_csr.py:108-151, launched by
 * _run_sddmm_prepared :257-330).  x is n_rows x k_dim and y n_cols x k_dim,
 * row-major contiguous, of P->value_type (FS_F32 / FS_F64); row_ids[p] is the
 * row of entry p (int32); columns int32.  accumulate: FS_ACC_NATIVE sums in
 * the value type, FS_ACC_WIDE in fp64; out_type: the value type or FS_F64.
 */
int fs_sddmm_csr(const fs_csr *P, const int32_t *row_ids, const void *x, const void *y,
                 int64_t k_dim, void *out, int32_t out_type, int32_t accumulate, double alpha,
                 double beta, void *stream);

/* ------------------------------------------------ MatrixMarket input */
/*
 * Replaces the reference harness's Python loader load_mtx_to_csr_torch
 * (tests/test_spmv.py:103-176) with its reading rules: banner field /
 * symmetry, '%' comments, the first `entries` data lines, pattern = 1.0,
 * complex files read as their real part, entries outside the matrix dropped,
 * symmetric / hermitian mirrored, skew-symmetric mirrored negated, repeated
 * pairs summed in fp64 in file order, columns ascending per row.
 * fs_mtx_read parses into host memory and reports the CSR sizes;
 * fs_mtx_fill copies row_ptr[n_rows + 1], col_idx[nnz] (int64) and
 * values[nnz] (fp64) out (any of them may be null); fs_mtx_free releases.
 */
enum { FS_MTX_REAL = 0, FS_MTX_INTEGER = 1, FS_MTX_COMPLEX = 2, FS_MTX_PATTERN = 3 };
enum { FS_MTX_GENERAL = 0, FS_MTX_SYMMETRIC = 1, FS_MTX_SKEW = 2, FS_MTX_HERMITIAN = 3 };

typedef struct fs_mtx_info {
  int64_t n_rows;
  int64_t n_cols;
  int64_t nnz;        /* after mirroring and summing repeated pairs */
  int32_t field;      /* FS_MTX_REAL ... FS_MTX_PATTERN            */
  int32_t symmetry;   /* FS_MTX_GENERAL ... FS_MTX_HERMITIAN       */
} fs_mtx_info;

int fs_mtx_read(const char *path, fs_mtx_info *info, void **handle);
int fs_mtx_fill(void *handle, int64_t *row_ptr, int64_t *col_idx, double *values);
void fs_mtx_free(void *handle);
const char *fs_mtx_last_error(void);

/* ------------------------------------------------- column-sliced SpMV */
/*
 * No counterpart in the reference (whose SpMV gathers x in place,
 * spmv_csr.py:158-184); used by the fp32 / fp64 / complex native path when x
 * is larger than one XCD's L2.  The matrix is stored column-slice-major:
 * every (slice, row) pair with entries is a compact row, ordered by slice,
 * then row.  fs_spmv_csr over the compact rows (absolute column indices)
 * writes one partial per compact row into z (as y, or as raw fp64 row sums
 * through fs_plan.row_sums); fs_sliced_combine then forms
 * y[r] = alpha * sum over slices of z[compact row (k, r)] + beta * y[r],
 * summed in slice order in fp64 (or compensated fp32, fs_slices.fp32_sums).
 */
typedef struct fs_slices {
  int32_t value_type;        /* of y: FS_F32, FS_F64, FS_C64 or FS_C128    */
  int32_t n_slices;
  int64_t n_rows;            /* rows of the original matrix                */
  int64_t n_compact;
  const int32_t *compact_row;  /* [n_compact] row of each compact row      */
  const int64_t *bounds;     /* [n_slices * (n_blocks + 1)], n_blocks =
                                ceil(n_rows / rows_per_block): first compact
                                row of slice k in row block b               */
  int32_t rows_per_block;    /* 256                                        */
  int32_t z_type;            /* of the partials: value_type, or FS_F64 /
                                FS_C128 for row sums (fs_plan.row_sums)     */
  const uint32_t *row_mask;  /* null, or [n_slices * n_blocks * 8] words
                                (rows_per_block 256): bit r % 32 of word
                                (k * n_blocks + b) * 8 + (r % 256) / 32 is
                                set when row r = 256 b + (r % 256) has a
                                compact row in slice k.  With it, a row's
                                partial in slice k is found by counting the
                                set bits before it, and compact_row is not
                                read.                                       */
  int32_t fp32_sums;         /* 1: add fp32 partials of fp32 values in
                                compensated fp32 (the split-precision
                                modes); 0: in fp64                          */
  int32_t reserved;
} fs_slices;

int fs_sliced_combine(const fs_slices *sl, const void *z, void *y, double alpha, double beta,
                      void *stream);

/* ---------------------------------------------------- x-window SpMV */
/*
 * No counterpart in the reference (whose SpMV gathers x in place,
 * spmv_csr.py:158-184).  First pass of a two-pass SpMV for x far larger than
 * the L2: the columns are cut into windows of `window` columns, and a second
 * copy of the column indices is stored window-major (window by window; inside
 * a window by entry tile, then in CSR order) as 16-bit in-window columns.  A
 * run is the entries of one entry tile inside one window.  fs_xwin_gather
 * stages each window of x in one workgroup's LDS and writes every entry's x
 * value to xg: entry k (window-major position) of run r goes to
 * xg[k + run_off[r]] (tile-major xg), or to xg[k] (window-major xg).  The
 * second pass is fs_spmv_csr with entry tiles and fs_plan.x_perm (and, for a
 * window-major xg, fs_plan.xg_run_off).
 */
typedef struct fs_xwin {
  int32_t value_type;           /* of x and xg: FS_F32 or FS_F64                */
  int32_t window;               /* columns per window: a multiple of 4, at most
                                   fs_xwin_max_window(value_type)               */
  int64_t n_cols;
  int64_t nnz;
  int64_t n_runs;
  int64_t n_parts;              /* workgroups of the gather                     */
  const uint16_t *wcol;         /* [nnz] column - window * (column / window)    */
  const int32_t *batch_run;     /* [ceil(nnz / 64)] run of entry 64 b           */
  const uint64_t *batch_mask;   /* [ceil(nnz / 64)] bit i (i >= 1) set when
                                   entry 64 b + i starts a run                  */
  const int32_t *run_off;       /* [n_runs]                                     */
  const int32_t *part_window;   /* [n_parts] window holding part p's first entry */
  const int64_t *part_entry;    /* [n_parts + 1] part p = window-major entries
                                   [part_entry[p], part_entry[p + 1]); a part
                                   may run through several windows             */
  const int64_t *window_entry;  /* [ceil(n_cols / window) + 1] first
                                   window-major entry of each window          */
  int32_t order;                /* FS_XG_TILE_MAJOR: entry k goes to
                                   xg[k + run_off[r]] (tile order, read by the
                                   second pass as one stream per tile);
                                   FS_XG_WINDOW_MAJOR: entry k goes to xg[k]
                                   (batch_run / batch_mask / run_off unused; the
                                   second pass finds its values through
                                   fs_plan.xg_run_off; wcol and xg 16-byte
                                   aligned)                                    */
  int32_t reserved;
} fs_xwin;

enum { FS_XG_TILE_MAJOR = 0, FS_XG_WINDOW_MAJOR = 1 };

/* Largest window (columns) one workgroup can stage on the current device, or -1. */
int64_t fs_xwin_max_window(int32_t value_type);
int fs_xwin_gather(const fs_xwin *L, const void *x, void *xg, void *stream);

/* Bytes of `partials` needed for n_chunks chunks (n_dense = 1 for SpMV). */
int64_t fs_partials_bytes(int32_t value_type, int32_t accumulate, int64_t n_chunks,
                          int64_t n_dense);

/*
 * y = alpha * A x + beta * y.  y is not read when beta == 0.  x has n_cols
 * entries and y n_rows entries, both of A->value_type.
 */
int fs_spmv_csr(const fs_csr *A, const void *x, void *y, double alpha, double beta,
                const fs_plan *plan, void *stream);

/*
 * fs_spmv_csr without a caller-owned partials buffer: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * plan->partials and plan->partials_bytes are ignored; the chunk partials /
 * tile carries live in a library-owned device array of
 * FS_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.  A plan that
 * needs more (fs_tile_carry_bytes / fs_partials_bytes) is rejected with
 * FS_ERR_UNSUPPORTED before anything is launched.  The call allocates
 * nothing.  Calls that overlap on one GPU (several streams or host threads)
 * share the array and race; order them on one stream.
 */
#define FS_NOALLOC_SCRATCH_BYTES (4u << 20)
int fs_spmv_csr_noalloc(const fs_csr *A, const void *x, void *y, double alpha, double beta,
                        const fs_plan *plan, void *stream);

/*
 * C = alpha * A B + beta * C with B (n_cols x n_dense, leading dimension ldb)
 * and C (n_rows x n_dense, leading dimension ldc), both row-major.  C is not
 * read when beta == 0.  col_tile is the number of dense columns one lane
 * group covers per pass (0 = automatic).
 */
int fs_spmm_csr(const fs_csr *A, const void *B, int64_t ldb, void *C, int64_t ldc,
                int64_t n_dense, double alpha, double beta, const fs_plan *plan,
                int32_t col_tile, void *stream);

/*
 * One device pass over row_ptr and col_idx.  Writes FS_STATS_LEN int64 values
 * to stats (device memory):
 *   [0] row_ptr[0]            [1] row_ptr[n_rows]
 *   [2] #i with row_ptr[i+1] < row_ptr[i]
 *   [3] min column index (INT64_MAX if nnz == 0)
 *   [4] max column index (-1 if nnz == 0)
 *   [5] max row length        [6] min row length (0 if n_rows == 0)
 *   [7] sum of squared row lengths
 *   [8] empty rows
 *   [9..14] rows with length in (0,4], (4,16], (16,64], (64,256],
 *           (256,4096], (4096,inf)
 *   [15] entries in rows longer than 256
 * Column statistics are taken over the first nnz entries of col_idx; the row
 * statistics over the n_rows + 1 row pointers.
 */
#define FS_STATS_LEN 16
int fs_csr_analyze(const fs_csr *A, int64_t *stats, void *stream);

/* ------------------------------------------- split-precision value storage */
/*
 * Storage formats for the values of an fp32 matrix (x, y stay fp32; the sum is
 * fp32 with compensated summation in the chunked long-row role).  No
 * counterpart in the reference, whose values are always stored in the call's
 * dtype; used by callers that trade value precision for bytes.
 */
enum {
  FS_STORE_BF16X2 = 1,       /* hi = bf16(v), lo = bf16(v - hi); a = hi + lo  */
  FS_STORE_BF16 = 2,         /* hi plane only                                 */
  FS_STORE_F16_ROWSCALE = 3  /* h = fp16(v / s_r), s_r = 2^ceil(log2 max|row|) */
};

/*
 * Encode A's fp32 values (A->values) into a storage format: hi / lo are nnz
 * uint16 planes (lo only for FS_STORE_BF16X2; FS_STORE_F16_ROWSCALE writes the
 * fp16 plane to hi), row_scale has n_rows floats (FS_STORE_F16_ROWSCALE only).
 */
int fs_split_values(int32_t format, const fs_csr *A, void *hi, void *lo, float *row_scale,
                    void *stream);

/*
 * y = alpha * A x + beta * y with A's values taken from a storage format
 * written by fs_split_values (A->values is not read).  A must be FS_F32 with
 * FS_I32 column indices.
 */
int fs_spmv_csr_split(const fs_csr *A, int32_t format, const void *hi, const void *lo,
                      const float *row_scale, const void *x, void *y, double alpha, double beta,
                      const fs_plan *plan, void *stream);

/* ------------------------------------------------- batched persistent path */
/*
 * Many independent fp32 / int32-column CSR problems in one persistent launch
 * (replayable from a captured graph).  No counterpart in the reference, which
 * launches per matrix and row bucket; used by callers with many small or
 * uneven matrices.
 */
enum { FS_BATCH_SPMV = 0, FS_BATCH_SPMM = 1 };

typedef struct fs_batch_entry {
  fs_csr A;              /* FS_F32 values, FS_I32 columns                     */
  const void *x;         /* SpMV operand (n_cols), or NULL                    */
  void *y;               /* SpMV result (n_rows)                              */
  const void *B;         /* SpMM operand (n_cols x n_dense, ldb), or NULL     */
  void *C;               /* SpMM result (n_rows x n_dense, ldc)               */
  int64_t n_dense, ldb, ldc;
  int32_t group_lanes;   /* SpMV lanes per row: 1..64, power of two           */
  int32_t reserved;
} fs_batch_entry;

/* One unit of work: rows [row_begin, row_end) of matrix `matrix`. */
typedef struct fs_batch_item {
  int32_t matrix;
  int32_t kind;          /* FS_BATCH_SPMV / FS_BATCH_SPMM */
  int64_t row_begin;
  int64_t row_end;
} fs_batch_item;

/* Bytes of one packed matrix record; fs_batch_pack writes n_entries records
 * into host memory, which the caller copies to the device. */
int64_t fs_batch_matrix_bytes(void);
int fs_batch_pack(const fs_batch_entry *entries, int64_t n_entries, void *packed_host);

/*
 * Run every item: y = alpha A x + beta y for SpMV items, C = alpha A B + beta C
 * for SpMM items.  packed_dev: the packed records in device memory; items:
 * device array; queue: two int32 in device memory, zero before the first call
 * and zero again after every call.  grid: persistent block count (e.g. a
 * small multiple of the CU count).  Single-row SpMV items longer than
 * long_row entries are reduced by a whole block.  Calls sharing a queue must
 * be ordered on one stream.
 */
int fs_batch_run(const void *packed_dev, int64_t n_matrices, const fs_batch_item *items,
                 int64_t n_items, int64_t long_row, int32_t *queue, int32_t grid, double alpha,
                 double beta, void *stream);

/* --------------------------------------------------- learned kernel choice */
/*
 * A 16 -> hidden -> hidden -> out_dim perceptron (ReLU hidden layers) whose
 * outputs are predicted relative times of out_dim schedule variants; the
 * selector picks the smallest.  params is one fp32 array laid out as
 * W1[in][hidden], b1[hidden], W2[hidden][hidden], b2[hidden],
 * W3[hidden][out], b3[out] (row-major).  No counterpart in the reference,
 * which picks launch parameters from static tables (spmv_csr.py:128-153,
 * spmm_csr_opt_alg2.py:175-304).
 */
typedef struct fs_mlp {
  int32_t in_dim;
  int32_t hidden;
  int32_t out_dim;
  int32_t reserved;
  float *params;
} fs_mlp;

int64_t fs_mlp_param_count(int32_t in_dim, int32_t hidden, int32_t out_dim);

/* 16 float features of a matrix from the FS_STATS_LEN statistics of
 * fs_csr_analyze (device pointers): log2(n_rows+1), log2(n_cols+1),
 * log2(nnz+1), mean, std, std/mean, log2(max+1), min, empty fraction, six
 * length-bin fractions, fraction of entries in rows longer than 256. */
int fs_csr_features(const int64_t *stats, int64_t n_rows, int64_t n_cols, int64_t nnz,
                    float *features, void *stream);

/* Forward pass over batch samples X[batch][in]; H1, H2 ([batch][hidden]) may be
 * null when no backward pass follows; Y is [batch][out]. */
int fs_mlp_forward(const fs_mlp *net, const float *X, int64_t batch, float *H1, float *H2,
                   float *Y, void *stream);

/* Gradient (same layout as params) of loss = sum(0.5 (M (Y - T))^2) /
 * (batch * out) and the loss itself (one float, may be null).  workspace of
 * fs_mlp_workspace_bytes bytes.  Deterministic. */
int64_t fs_mlp_workspace_bytes(int32_t in_dim, int32_t hidden, int32_t out_dim, int64_t batch);
int fs_mlp_backward(const fs_mlp *net, const float *X, const float *H1, const float *H2,
                    const float *Y, const float *T, const float *M, int64_t batch, float *grad,
                    float *loss, void *workspace, int64_t workspace_bytes, void *stream);

/* Bias-corrected Adam step (step counts from 1) over n parameters. */
int fs_adam_step(float *params, const float *grad, float *m, float *v, int64_t n, double lr,
                 double beta1, double beta2, double eps, int64_t step, void *stream);

/* Standardise features ((f - mu) / sd), forward them and write the index of the
 * smallest output to *choice (and all outputs to scores unless null). */
int fs_mlp_select(const fs_mlp *net, const float *features, const float *mu, const float *sd,
                  int32_t *choice, float *scores, void *stream);

/* Message of the calling thread's last failed call ("" if none). */
const char *fs_last_error(void);

/* Library version string. */
const char *fs_version(void);

#ifdef __cplusplus
}
#endif

#endif /* FLAGSPARSE_HIP_H */
