// batched.hip -- many independent fp32 CSR SpMV / SpMM problems in one
// persistent launch.
//
// The work list (device memory, built once per batch) holds items of
// (matrix, kind, row range); ranges are cut so every item carries about the
// same number of entries (weighted by dense columns for SpMM).  A persistent
// grid of blocks pulls items from a device work queue (one atomic per item).
// Rows are computed with the same per-row procedures (lane-group width,
// unroll, reduction tree) as the single-matrix kernels' short-row role, so a
// row's result is bit-identical to the single-matrix launch with the same
// group width.  A single-row item longer than a lane group's bound is reduced
// by the whole block instead.
//
// The queue head and the finished-block counter reset themselves when the last
// block leaves, so one launch can be replayed from a captured graph with no
// memset node.  No counterpart in the reference (one launch per matrix and
// bucket there).
#include "spmv_kernel.hpp"

namespace fs {

constexpr int kBatBlock = 256;

struct BatchMatrix {
  const void* row_ptr;
  const int32_t* col;
  const float* val;
  const float* x;
  float* y;
  const float* B;
  float* C;
  int64_t n_rows, n_cols, nnz, n_dense, ldb, ldc;
  int32_t group_lanes, ptr64;
  int32_t mm_lanes, mm_vw;  // SpMM lane group and vector width
};

__device__ __forceinline__ int64_t bm_ptr(const BatchMatrix& m, int64_t r) {
  return m.ptr64 ? static_cast<const int64_t*>(m.row_ptr)[r]
                 : (int64_t) static_cast<const int32_t*>(m.row_ptr)[r];
}

// --- SpMV rows of an item with groups of S lanes (same procedure as the
// single-matrix short role: accumulate_span<S, 4> then group_reduce<S>).
template <int S>
__device__ void bat_spmv_rows(const BatchMatrix& m, int64_t r0, int64_t r1, double alpha, double beta) {
  constexpr int kGroups = kBatBlock / S;
  const int grp = threadIdx.x / S, g = threadIdx.x % S;
  DirectVals<float> vs{m.val};
  // every group runs the same number of iterations so the S-lane exchange in
  // group_reduce always sees its whole group active
  for (int64_t base = r0; base < r1; base += kGroups) {
    const int64_t row = base + grp;
    const bool live = row < r1;
    Plain<float> acc;
    acc.init();
    if (live) {
      const int64_t rs = bm_ptr(m, row), re = bm_ptr(m, row + 1);
      accumulate_span<S, 4>(acc, m.col, vs, m.x, rs, re, g);
    }
    group_reduce<S>(acc);
    if (live && g == 0) write_out(m.y, row, acc.value(), alpha, beta);
  }
}

// --- one long SpMV row with the whole block: lanes stride by 256, wave
// reduction, then the 4 wave totals summed in order through LDS.
__device__ void bat_spmv_long_row(const BatchMatrix& m, int64_t row, double alpha, double beta,
                                  float* lds) {
  DirectVals<float> vs{m.val};
  Comp acc;
  acc.init();
  const int64_t rs = bm_ptr(m, row), re = bm_ptr(m, row + 1);
  accumulate_span<kBatBlock, 4>(acc, m.col, vs, m.x, rs, re, threadIdx.x);
  group_reduce<kWave>(acc);
  const int w = threadIdx.x / kWave;
  if (threadIdx.x % kWave == 0) { lds[2 * w] = acc.s; lds[2 * w + 1] = acc.c; }
  __syncthreads();
  if (threadIdx.x == 0) {
    Comp t;
    t.init();
    for (int i = 0; i < kBatBlock / kWave; ++i) {
      Comp o;
      o.s = lds[2 * i];
      o.c = lds[2 * i + 1];
      t.merge(o);
    }
    write_out(m.y, row, t.value(), alpha, beta);
  }
  __syncthreads();
}

// --- SpMM rows: groups of G lanes, VW columns per lane (the single-matrix
// short role's mm_accumulate<VW, 4>, restated here for the batch layout).
constexpr int kBatMmU = 8;  // SpMM entries in flight per lane (4: 12.0 ms on the 512-matrix batch, 8: 11.0 ms)

template <int G, int VW>
__device__ void bat_spmm_rows(const BatchMatrix& m, int64_t r0, int64_t r1, double alpha, double beta) {
  constexpr int kGroups = kBatBlock / G;
  const int grp = threadIdx.x / G, g = threadIdx.x % G;
  const int64_t tiles = (m.n_dense + G * VW - 1) / (G * VW);
  for (int64_t t = 0; t < tiles; ++t) {
    const int64_t n0 = t * (G * VW) + (int64_t)g * VW;
    if (n0 >= m.n_dense) continue;
    for (int64_t row = r0 + grp; row < r1; row += kGroups) {
      const int64_t rs = bm_ptr(m, row), re = bm_ptr(m, row + 1);
      float acc[VW];
#pragma unroll
      for (int k = 0; k < VW; ++k) acc[k] = 0.f;
      int64_t j = rs;
      for (; j + kBatMmU - 1 < re; j += kBatMmU) {
        int64_t col[kBatMmU];
        float a[kBatMmU];
#pragma unroll
        for (int u = 0; u < kBatMmU; ++u) { col[u] = m.col[j + u]; a[u] = m.val[j + u]; }
        float b[kBatMmU][VW];
#pragma unroll
        for (int u = 0; u < kBatMmU; ++u) {
          if constexpr (VW == 4) {
            const float4 v = *reinterpret_cast<const float4*>(m.B + col[u] * m.ldb + n0);
            b[u][0] = v.x; b[u][1] = v.y; b[u][2] = v.z; b[u][3] = v.w;
          } else {
#pragma unroll
            for (int k = 0; k < VW; ++k) b[u][k] = m.B[col[u] * m.ldb + n0 + k];
          }
        }
#pragma unroll
        for (int u = 0; u < kBatMmU; ++u)
#pragma unroll
          for (int k = 0; k < VW; ++k) acc[k] = __builtin_fmaf(a[u], b[u][k], acc[k]);
      }
      if (j < re) {  // the last, partial batch: its loads in flight together
        int32_t ci[kBatMmU];
        float a[kBatMmU];
#pragma unroll
        for (int u = 0; u < kBatMmU; ++u) {
          const bool ok = j + u < re;
          ci[u] = ok ? m.col[j + u] : 0;
          a[u] = ok ? m.val[j + u] : 0.f;
        }
        // the empty asm keeps each index's widening (and the wait for its
        // load) out of the masked load's branch, so all four loads are issued
        // before the first wait
        int64_t col[kBatMmU];
#pragma unroll
        for (int u = 0; u < kBatMmU; ++u) {
          asm volatile("" : "+v"(ci[u]));
          col[u] = ci[u];
        }
        float b[kBatMmU][VW];
#pragma unroll
        for (int u = 0; u < kBatMmU; ++u) {
#pragma unroll
          for (int k = 0; k < VW; ++k) b[u][k] = 0.f;
          if (j + u < re) {
            if constexpr (VW == 4) {
              const float4 v = *reinterpret_cast<const float4*>(m.B + col[u] * m.ldb + n0);
              b[u][0] = v.x; b[u][1] = v.y; b[u][2] = v.z; b[u][3] = v.w;
            } else {
#pragma unroll
              for (int k = 0; k < VW; ++k) b[u][k] = m.B[col[u] * m.ldb + n0 + k];
            }
          }
        }
#pragma unroll
        for (int u = 0; u < kBatMmU; ++u)
          if (j + u < re)
#pragma unroll
            for (int k = 0; k < VW; ++k) acc[k] = __builtin_fmaf(a[u], b[u][k], acc[k]);
      }
      if constexpr (VW == 4) {  // write_out's arithmetic, one 16-byte store
        float* c = m.C + row * m.ldc + n0;
        float v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = alpha == 1.0 ? acc[k] : scale(acc[k], alpha);
        if (beta != 0.0) {
          const float4 o = *reinterpret_cast<const float4*>(c);
          v[0] = add(v[0], scale(o.x, beta));
          v[1] = add(v[1], scale(o.y, beta));
          v[2] = add(v[2], scale(o.z, beta));
          v[3] = add(v[3], scale(o.w, beta));
        }
        *reinterpret_cast<float4*>(c) = make_float4(v[0], v[1], v[2], v[3]);
      } else {
#pragma unroll
        for (int k = 0; k < VW; ++k) write_out(m.C, row * m.ldc + n0 + k, acc[k], alpha, beta);
      }
    }
  }
}

__device__ void bat_spmv_dispatch(const BatchMatrix& m, int64_t r0, int64_t r1, double alpha, double beta) {
  switch (m.group_lanes) {
    case 1: bat_spmv_rows<1>(m, r0, r1, alpha, beta); break;
    case 2: bat_spmv_rows<2>(m, r0, r1, alpha, beta); break;
    case 4: bat_spmv_rows<4>(m, r0, r1, alpha, beta); break;
    case 8: bat_spmv_rows<8>(m, r0, r1, alpha, beta); break;
    case 16: bat_spmv_rows<16>(m, r0, r1, alpha, beta); break;
    case 32: bat_spmv_rows<32>(m, r0, r1, alpha, beta); break;
    default: bat_spmv_rows<64>(m, r0, r1, alpha, beta); break;
  }
}

__device__ void bat_spmm_dispatch(const BatchMatrix& m, int64_t r0, int64_t r1, double alpha, double beta) {
  if (m.mm_vw == 4) {
    switch (m.mm_lanes) {
      case 1: bat_spmm_rows<1, 4>(m, r0, r1, alpha, beta); break;
      case 2: bat_spmm_rows<2, 4>(m, r0, r1, alpha, beta); break;
      case 4: bat_spmm_rows<4, 4>(m, r0, r1, alpha, beta); break;
      case 8: bat_spmm_rows<8, 4>(m, r0, r1, alpha, beta); break;
      case 16: bat_spmm_rows<16, 4>(m, r0, r1, alpha, beta); break;
      case 32: bat_spmm_rows<32, 4>(m, r0, r1, alpha, beta); break;
      default: bat_spmm_rows<64, 4>(m, r0, r1, alpha, beta); break;
    }
  } else {
    switch (m.mm_lanes) {
      case 1: bat_spmm_rows<1, 1>(m, r0, r1, alpha, beta); break;
      case 2: bat_spmm_rows<2, 1>(m, r0, r1, alpha, beta); break;
      case 4: bat_spmm_rows<4, 1>(m, r0, r1, alpha, beta); break;
      case 8: bat_spmm_rows<8, 1>(m, r0, r1, alpha, beta); break;
      case 16: bat_spmm_rows<16, 1>(m, r0, r1, alpha, beta); break;
      case 32: bat_spmm_rows<32, 1>(m, r0, r1, alpha, beta); break;
      default: bat_spmm_rows<64, 1>(m, r0, r1, alpha, beta); break;
    }
  }
}

__global__ __launch_bounds__(kBatBlock) void batched_kernel(const BatchMatrix* __restrict__ mats,
                                                            const fs_batch_item* __restrict__ items,
                                                            int64_t n_items, int64_t long_row,
                                                            double alpha, double beta,
                                                            int32_t* __restrict__ queue) {
  __shared__ int64_t cur;
  __shared__ float lds[2 * (kBatBlock / kWave)];
  while (true) {
    if (threadIdx.x == 0)
      cur = __hip_atomic_fetch_add(queue, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int64_t it = cur;
    __syncthreads();
    if (it >= n_items) break;
    const fs_batch_item w = items[it];
    const BatchMatrix& m = mats[w.matrix];
    if (w.kind == FS_BATCH_SPMV) {
      if (w.row_end - w.row_begin == 1 &&
          bm_ptr(m, w.row_begin + 1) - bm_ptr(m, w.row_begin) > long_row)
        bat_spmv_long_row(m, w.row_begin, alpha, beta, lds);
      else
        bat_spmv_dispatch(m, w.row_begin, w.row_end, alpha, beta);
    } else {
      bat_spmm_dispatch(m, w.row_begin, w.row_end, alpha, beta);
    }
  }
  // the last block out resets the queue for the next launch
  if (threadIdx.x == 0) {
    const int done = __hip_atomic_fetch_add(queue + 1, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (done == (int)gridDim.x - 1) {
      __hip_atomic_store(queue, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(queue + 1, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

static int pow2_ceil_i(int64_t v) {
  int g = 1;
  while (g < v && g < 64) g <<= 1;
  return g;
}

}  // namespace fs

using namespace fs;

extern "C" int64_t fs_batch_matrix_bytes(void) { return (int64_t)sizeof(BatchMatrix); }

extern "C" int fs_batch_pack(const fs_batch_entry* entries, int64_t n_entries, void* packed_host) {
  if (n_entries < 0 || (n_entries > 0 && (!entries || !packed_host)))
    return fail(FS_ERR_VALUE, "fs_batch_pack: bad arguments");
  BatchMatrix* out = static_cast<BatchMatrix*>(packed_host);
  for (int64_t i = 0; i < n_entries; ++i) {
    const fs_batch_entry& e = entries[i];
    const fs_csr& A = e.A;
    int rc = check_csr(&A, "fs_batch_pack");
    if (rc) return rc;
    if (A.value_type != FS_F32 || A.index_type != FS_I32)
      return fail(FS_ERR_TYPE, "fs_batch_pack: entry %lld must be fp32 with int32 columns", (long long)i);
    if (e.n_dense < 0 || (e.n_dense > 0 && (e.ldb < e.n_dense || e.ldc < e.n_dense)))
      return fail(FS_ERR_VALUE, "fs_batch_pack: entry %lld has bad dense shape", (long long)i);
    const int s = e.group_lanes;
    if (s != 1 && s != 2 && s != 4 && s != 8 && s != 16 && s != 32 && s != 64)
      return fail(FS_ERR_VALUE, "fs_batch_pack: entry %lld group_lanes must be a power of two <= 64", (long long)i);
    BatchMatrix m;
    m.row_ptr = A.row_ptr;
    m.col = static_cast<const int32_t*>(A.col_idx);
    m.val = static_cast<const float*>(A.values);
    m.x = static_cast<const float*>(e.x);
    m.y = static_cast<float*>(e.y);
    m.B = static_cast<const float*>(e.B);
    m.C = static_cast<float*>(e.C);
    m.n_rows = A.n_rows;
    m.n_cols = A.n_cols;
    m.nnz = A.nnz;
    m.n_dense = e.n_dense;
    m.ldb = e.ldb;
    m.ldc = e.ldc;
    m.group_lanes = s;
    m.ptr64 = A.ptr_type == FS_I64;
    int vw = 1;
    if (e.n_dense > 0 && e.n_dense % 4 == 0 && e.ldb % 4 == 0 && e.ldc % 4 == 0 &&
        ((uintptr_t)e.B % 16) == 0 && ((uintptr_t)e.C % 16) == 0)
      vw = 4;
    m.mm_vw = vw;
    m.mm_lanes = pow2_ceil_i((e.n_dense + vw - 1) / vw);
    out[i] = m;
  }
  return FS_OK;
}

extern "C" int fs_batch_run(const void* packed_dev, int64_t n_matrices, const fs_batch_item* items,
                            int64_t n_items, int64_t long_row, int32_t* queue, int32_t grid,
                            double alpha, double beta, void* stream) {
  if (n_items < 0 || n_matrices < 0) return fail(FS_ERR_VALUE, "fs_batch_run: negative counts");
  if (n_items == 0) return FS_OK;
  if (!packed_dev || !items || !queue) return fail(FS_ERR_VALUE, "fs_batch_run: null pointer");
  if (grid <= 0) return fail(FS_ERR_VALUE, "fs_batch_run: grid must be positive");
  if (long_row < 1) return fail(FS_ERR_VALUE, "fs_batch_run: long_row must be >= 1");
  // the work queue head is an int32 that every block advances once past the
  // last item; it must not wrap
  if (n_items + (int64_t)grid > (int64_t)INT32_MAX)
    return fail(FS_ERR_UNSUPPORTED, "fs_batch_run: n_items + grid must stay below 2^31 - 1");
  batched_kernel<<<grid, kBatBlock, 0, as_stream(stream)>>>(static_cast<const BatchMatrix*>(packed_dev),
                                                           items, n_items, long_row, alpha, beta, queue);
  return check_launch("fs_batch_run");
}
