// analyze.hip -- one device pass that replaces the reference's host-synchronising
// validation reductions (spmv_csr.py:525-542, spmm_csr.py:226-252: indptr[0],
// indptr[-1], monotonicity, min/max column, max row length) and gathers the
// row-length statistics the kernel selector reads.  The caller copies the
// FS_STATS_LEN results to the host once.
//
// All statistics are integers, combined with integer atomics, so the result is
// exact and independent of scheduling.
#include "fs_host.hpp"

namespace fs {

typedef int32_t i32x4_t __attribute__((ext_vector_type(4)));

constexpr int kAnBlock = 1024;

enum {
  ST_PTR0 = 0, ST_PTRN = 1, ST_DEC = 2, ST_CMIN = 3, ST_CMAX = 4, ST_LMAX = 5, ST_LMIN = 6,
  ST_SUMSQ = 7, ST_EMPTY = 8, ST_HIST = 9, ST_LONGNNZ = 15
};
static_assert(ST_LONGNNZ + 1 == FS_STATS_LEN, "statistics layout");

template <typename P>
__global__ void analyze_init(const P* ptr, int64_t n_rows, int64_t* st) {
  const int t = threadIdx.x;
  if (t < FS_STATS_LEN) {
    int64_t v = 0;
    if (t == ST_CMIN) v = INT64_MAX;
    else if (t == ST_CMAX) v = -1;
    else if (t == ST_LMIN) v = n_rows > 0 ? INT64_MAX : 0;
    else if (t == ST_PTR0) v = ptr ? (int64_t)ptr[0] : 0;
    else if (t == ST_PTRN) v = ptr ? (int64_t)ptr[n_rows] : 0;
    st[t] = v;
  }
}

__device__ __forceinline__ int64_t wave_sum(int64_t v) {
  for (int k = 32; k >= 1; k >>= 1) v += __shfl_xor(v, k);
  return v;
}
__device__ __forceinline__ int64_t wave_max(int64_t v) {
  for (int k = 32; k >= 1; k >>= 1) { int64_t o = __shfl_xor(v, k); v = o > v ? o : v; }
  return v;
}
__device__ __forceinline__ int64_t wave_min(int64_t v) {
  for (int k = 32; k >= 1; k >>= 1) { int64_t o = __shfl_xor(v, k); v = o < v ? o : v; }
  return v;
}

// Reduction of statistic k: 0 sum, 1 max, 2 min (PTR0 / PTRN are set by
// analyze_init and not reduced).
__device__ __forceinline__ int stat_op(int k) {
  return (k == ST_LMAX || k == ST_CMAX) ? 1 : (k == ST_LMIN || k == ST_CMIN) ? 2 : 0;
}

// Reduces the block's statistics v[2..FS_STATS_LEN) together: wave reductions,
// one LDS exchange and one barrier, then thread k applies statistic k with
// one atomic.
__device__ __forceinline__ void block_apply_all(int64_t (&v)[FS_STATS_LEN], int64_t* st) {
  __shared__ int64_t red[FS_STATS_LEN][kAnBlock / kWave];
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
#pragma unroll
  for (int k = 2; k < FS_STATS_LEN; ++k) {
    const int op = stat_op(k);
    v[k] = op == 0 ? wave_sum(v[k]) : op == 1 ? wave_max(v[k]) : wave_min(v[k]);
    if (lane == 0) red[k][w] = v[k];
  }
  __syncthreads();
  const int k = threadIdx.x;
  if (k >= 2 && k < FS_STATS_LEN) {
    const int op = stat_op(k);
    int64_t r = red[k][0];
    for (int i = 1; i < kAnBlock / kWave; ++i) {
      const int64_t o = red[k][i];
      r = op == 0 ? r + o : op == 1 ? (o > r ? o : r) : (o < r ? o : r);
    }
    int64_t* dst = st + k;
    if (op == 0) { if (r) __hip_atomic_fetch_add(dst, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
    else if (op == 1) __hip_atomic_fetch_max(dst, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else __hip_atomic_fetch_min(dst, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <typename I, typename P>
__global__ __launch_bounds__(kAnBlock) void analyze_pass(const P* __restrict__ ptr,
                                                          const I* __restrict__ idx,
                                                          int64_t n_rows, int64_t nnz,
                                                          int64_t* st) {
  const int64_t stride = (int64_t)gridDim.x * kAnBlock;
  const int64_t t0 = (int64_t)blockIdx.x * kAnBlock + threadIdx.x;
  int64_t dec = 0, lmax = 0, lmin = INT64_MAX, sumsq = 0, empty = 0, longnnz = 0;
  int64_t h[6] = {0, 0, 0, 0, 0, 0};
  // rows: four row pointer pairs in flight; lanes past the end re-read row
  // i0 and skip it
  constexpr int kRows = 4;
  for (int64_t i0 = t0; i0 < n_rows; i0 += kRows * stride) {
    int64_t len[kRows];
#pragma unroll
    for (int u = 0; u < kRows; ++u) {
      const int64_t i = i0 + u * stride < n_rows ? i0 + u * stride : i0;
      len[u] = (int64_t)ptr[i + 1] - (int64_t)ptr[i];
    }
#pragma unroll
    for (int u = 0; u < kRows; ++u) {
      if (i0 + u * stride >= n_rows) break;
      const int64_t l = len[u];
      dec += l < 0;
      lmax = l > lmax ? l : lmax;
      lmin = l < lmin ? l : lmin;
      sumsq += l * l;
      if (l <= 0) empty += l == 0;
      else if (l <= 4) ++h[0];
      else if (l <= 16) ++h[1];
      else if (l <= 64) ++h[2];
      else if (l <= 256) ++h[3];
      else if (l <= 4096) ++h[4];
      else ++h[5];
      if (l > 256) longnnz += l;
    }
  }
  // columns: only min and max are kept, so a lane past the end re-reads the
  // iteration's first column instead of branching
  int64_t cmin = INT64_MAX, cmax = -1;
  int64_t j0 = 0;
  if constexpr (sizeof(I) == 4) {
    // 16-byte loads (four columns), eight in flight per thread, over the
    // aligned body of the column array.  Non-temporal: the columns are read
    // once here, and an unprepared call's SpMV follows at once, whose data
    // they would otherwise push out of the caches (4M-row power-law matrix:
    // this pass 51.5 -> 53.8 us, the whole repeated unprepared call 268 ->
    // 254 us).
    const int64_t head = (int64_t)(((16 - ((uintptr_t)idx & 15)) & 15) / 4);
    const int64_t h0 = head < nnz ? head : nnz;
    const int64_t n4 = (nnz - h0) / 4;
    const int4* q = reinterpret_cast<const int4*>(idx + h0);
    int32_t mn = INT32_MAX, mx = INT32_MIN;
    for (int64_t k = t0; k < n4; k += 8 * stride) {
      int4 c[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const i32x4_t v = __builtin_nontemporal_load(
            reinterpret_cast<const i32x4_t*>(q + (k + u * stride < n4 ? k + u * stride : k)));
        c[u] = make_int4(v.x, v.y, v.z, v.w);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        mn = min(mn, min(min(c[u].x, c[u].y), min(c[u].z, c[u].w)));
        mx = max(mx, max(max(c[u].x, c[u].y), max(c[u].z, c[u].w)));
      }
    }
    if (t0 < n4) {
      cmin = mn;
      cmax = mx;
    }
    // the unaligned head [0, h0); the tail [h0 + 4 n4, nnz) is left to the
    // loop below
    if (t0 < h0) {
      const int64_t c = (int64_t)idx[t0];
      cmin = c < cmin ? c : cmin;
      cmax = c > cmax ? c : cmax;
    }
    j0 = h0 + 4 * n4;
  }
  for (int64_t j = j0 + t0; j < nnz; j += 4 * stride) {
    int64_t c[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) c[u] = (int64_t)idx[j + u * stride < nnz ? j + u * stride : j];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      cmin = c[u] < cmin ? c[u] : cmin;
      cmax = c[u] > cmax ? c[u] : cmax;
    }
  }
  int64_t v[FS_STATS_LEN] = {0, 0, dec, cmin, cmax, lmax, lmin, sumsq, empty,
                             h[0], h[1], h[2], h[3], h[4], h[5], longnnz};
  block_apply_all(v, st);
}

template <typename I, typename P>
int launch_analyze(const fs_csr* A, int64_t* st, hipStream_t s) {
  const P* ptr = static_cast<const P*>(A->row_ptr);
  const I* idx = static_cast<const I*>(A->col_idx);
  analyze_init<P><<<1, 64, 0, s>>>(ptr, A->n_rows, st);
  int rc = check_launch("fs_csr_analyze");
  if (rc) return rc;
  const int64_t work = A->n_rows > A->nnz ? A->n_rows : A->nnz;
  if (work == 0) return FS_OK;
  // few, wide blocks (two per CU): each reduces in LDS first, so every
  // statistic sees at most `blocks` same-address atomics (they serialise),
  // while each thread keeps eight 16-byte column loads in flight
  int64_t blocks = (work + kAnBlock * 32 - 1) / (kAnBlock * 32);
  if (blocks > 512) blocks = 512;
  analyze_pass<I, P><<<(unsigned)blocks, kAnBlock, 0, s>>>(ptr, idx, A->n_rows, A->nnz, st);
  return check_launch("fs_csr_analyze");
}

}  // namespace fs

extern "C" int fs_csr_analyze(const fs_csr* A, int64_t* stats, void* stream) {
  using namespace fs;
  if (!A) return fail(FS_ERR_VALUE, "fs_csr_analyze: matrix descriptor is null");
  if ((A->index_type != FS_I32 && A->index_type != FS_I64) ||
      (A->ptr_type != FS_I32 && A->ptr_type != FS_I64))
    return fail(FS_ERR_TYPE, "fs_csr_analyze: index and row-pointer types must be FS_I32 or FS_I64");
  if (A->n_rows < 0 || A->nnz < 0) return fail(FS_ERR_VALUE, "fs_csr_analyze: negative size");
  if (!stats) return fail(FS_ERR_VALUE, "fs_csr_analyze: stats is null");
  if (!A->row_ptr) return fail(FS_ERR_VALUE, "fs_csr_analyze: row_ptr is null");
  if (A->nnz > 0 && !A->col_idx) return fail(FS_ERR_VALUE, "fs_csr_analyze: col_idx is null");
  hipStream_t s = as_stream(stream);
  if (A->index_type == FS_I32 && A->ptr_type == FS_I32) return launch_analyze<int32_t, int32_t>(A, stats, s);
  if (A->index_type == FS_I32 && A->ptr_type == FS_I64) return launch_analyze<int32_t, int64_t>(A, stats, s);
  if (A->index_type == FS_I64 && A->ptr_type == FS_I32) return launch_analyze<int64_t, int32_t>(A, stats, s);
  return launch_analyze<int64_t, int64_t>(A, stats, s);
}
