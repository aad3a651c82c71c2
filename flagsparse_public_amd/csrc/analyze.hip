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

constexpr int kAnBlock = 256;

enum {
  ST_PTR0 = 0, ST_PTRN = 1, ST_DEC = 2, ST_CMIN = 3, ST_CMAX = 4, ST_LMAX = 5, ST_LMIN = 6,
  ST_SUMSQ = 7, ST_EMPTY = 8, ST_HIST = 9, ST_LONGNNZ = 15
};

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

// Reduces one value per thread over the block and applies it with one atomic.
template <int OP>  // 0 sum, 1 max, 2 min
__device__ __forceinline__ void block_apply(int64_t v, int64_t* dst, int64_t* lds) {
  v = OP == 0 ? wave_sum(v) : OP == 1 ? wave_max(v) : wave_min(v);
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t r = lds[0];
    for (int i = 1; i < kAnBlock / kWave; ++i) {
      const int64_t o = lds[i];
      r = OP == 0 ? r + o : OP == 1 ? (o > r ? o : r) : (o < r ? o : r);
    }
    if (OP == 0) { if (r) __hip_atomic_fetch_add(dst, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
    else if (OP == 1) __hip_atomic_fetch_max(dst, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else __hip_atomic_fetch_min(dst, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
}

template <typename I, typename P>
__global__ __launch_bounds__(kAnBlock) void analyze_pass(const P* __restrict__ ptr,
                                                          const I* __restrict__ idx,
                                                          int64_t n_rows, int64_t nnz,
                                                          int64_t* st) {
  __shared__ int64_t lds[kAnBlock / kWave];
  const int64_t stride = (int64_t)gridDim.x * kAnBlock;
  const int64_t t0 = (int64_t)blockIdx.x * kAnBlock + threadIdx.x;
  int64_t dec = 0, lmax = 0, lmin = INT64_MAX, sumsq = 0, empty = 0, longnnz = 0;
  int64_t h[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t i = t0; i < n_rows; i += stride) {
    const int64_t len = (int64_t)ptr[i + 1] - (int64_t)ptr[i];
    dec += len < 0;
    lmax = len > lmax ? len : lmax;
    lmin = len < lmin ? len : lmin;
    sumsq += len * len;
    if (len <= 0) empty += len == 0;
    else if (len <= 4) ++h[0];
    else if (len <= 16) ++h[1];
    else if (len <= 64) ++h[2];
    else if (len <= 256) ++h[3];
    else if (len <= 4096) ++h[4];
    else ++h[5];
    if (len > 256) longnnz += len;
  }
  int64_t cmin = INT64_MAX, cmax = -1;
  int64_t j = t0;
  for (; j + 3 * stride < nnz; j += 4 * stride) {  // four independent loads in flight
    int64_t c[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) c[u] = (int64_t)idx[j + u * stride];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      cmin = c[u] < cmin ? c[u] : cmin;
      cmax = c[u] > cmax ? c[u] : cmax;
    }
  }
  for (; j < nnz; j += stride) {
    const int64_t c = (int64_t)idx[j];
    cmin = c < cmin ? c : cmin;
    cmax = c > cmax ? c : cmax;
  }
  block_apply<0>(dec, st + ST_DEC, lds);
  block_apply<1>(lmax, st + ST_LMAX, lds);
  block_apply<2>(lmin, st + ST_LMIN, lds);
  block_apply<0>(sumsq, st + ST_SUMSQ, lds);
  block_apply<0>(empty, st + ST_EMPTY, lds);
  for (int b = 0; b < 6; ++b) block_apply<0>(h[b], st + ST_HIST + b, lds);
  block_apply<0>(longnnz, st + ST_LONGNNZ, lds);
  block_apply<2>(cmin, st + ST_CMIN, lds);
  block_apply<1>(cmax, st + ST_CMAX, lds);
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
  // few blocks: each reduces in LDS first, so every statistic sees at most
  // `blocks` same-address atomics (they serialise)
  int64_t blocks = (work + kAnBlock * 8 - 1) / (kAnBlock * 8);
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
