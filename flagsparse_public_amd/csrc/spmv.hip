// spmv.hip -- CSR SpMV for gfx950: y = alpha * A x + beta * y.
//
// One launch, two block roles (see fs_plan in flagsparse_hip.h):
//   * chunk blocks (lowest block ids, so the long work starts first): each
//     wavefront reduces one chunk of a listed long row; a row split over
//     several chunks is finished by whichever wavefront arrives last, which
//     sums the row's partials in chunk order -- the result does not depend on
//     arrival order.
//   * short blocks: groups of S lanes (S = 1..64) take consecutive rows in
//     natural order; each lane keeps U entries in flight (coalesced col/val
//     streams, U independent x gathers), then the group reduces with DPP /
//     swizzle exchanges.  Rows longer than short_max are skipped (chunked).
//
// Replaces the reference's row-per-program kernel (spmv_csr.py:158-184, the
// default path) and its bucketed CSR-vector kernels (spmv_csr.py:225-343,
// 393-440, the use_opt path) with a single wave64 schedule.
#include <type_traits>

#include "fs_host.hpp"

namespace fs {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;

template <typename V, typename I, typename P>
struct SpmvParams {
  const P* ptr;
  const I* idx;
  const V* val;
  const V* x;
  V* y;
  int64_t n_rows;
  double alpha, beta;
  int64_t short_max;
  int64_t n_chunk_blocks;
  int64_t n_chunks;
  int64_t chunk_nnz;
  const int64_t* long_rows;
  const int64_t* long_chunk_ptr;
  const int32_t* chunk_long;
  int32_t* counters;
  void* partials;
  int xcd_remap;
};

// Accumulator used for chunks: fp32-native chunked rows get compensation.
template <typename A> struct ChunkAcc { using T = A; };
template <> struct ChunkAcc<Plain<float>> { using T = Comp; };

// Coherent (device-scope) word-wise load/store of an accumulator, so chunk
// partials written by a wavefront on one XCD are read correctly on another.
template <typename A>
__device__ __forceinline__ void store_partial(A* p, const A& a) {
  static_assert(sizeof(A) % 4 == 0, "partials are stored as 32-bit words");
  const uint32_t* src = reinterpret_cast<const uint32_t*>(&a);
  uint32_t* dst = reinterpret_cast<uint32_t*>(p);
#pragma unroll
  for (int w = 0; w < (int)(sizeof(A) / 4); ++w)
    __hip_atomic_store(dst + w, src[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename A>
__device__ __forceinline__ A load_partial(const A* p) {
  A a;
  uint32_t* dst = reinterpret_cast<uint32_t*>(&a);
  const uint32_t* src = reinterpret_cast<const uint32_t*>(p);
#pragma unroll
  for (int w = 0; w < (int)(sizeof(A) / 4); ++w)
    dst[w] = __hip_atomic_load(src + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return a;
}

// Accumulate entries [s, e) of one row with `STRIDE` lanes, lane offset g,
// U entries in flight per lane.
template <int STRIDE, int U, typename A, typename V, typename I>
__device__ __forceinline__ void accumulate_span(A& acc, const I* __restrict__ idx,
                                                const V* __restrict__ val,
                                                const V* __restrict__ x, int64_t s,
                                                int64_t e, int g) {
  using C = typename A::compute_t;
  int64_t j = s + g;
  for (; j + (int64_t)(U - 1) * STRIDE < e; j += (int64_t)U * STRIDE) {
    int64_t col[U];
    C a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      col[u] = ldg_index(idx, j + u * STRIDE);
      a[u] = load_as<C>(val, j + u * STRIDE);
    }
    C xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xv[u] = load_as<C>(x, col[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) acc.add_prod(a[u], xv[u]);
  }
  for (; j < e; j += STRIDE) {
    const int64_t c = ldg_index(idx, j);
    acc.add_prod(load_as<C>(val, j), load_as<C>(x, c));
  }
}

template <typename V, typename A, typename I, typename P, int U>
__device__ __forceinline__ void spmv_chunk_role(const SpmvParams<V, I, P>& p, int64_t bid) {
  using CA = typename ChunkAcc<A>::T;
  const int lane = threadIdx.x % kWave;
  const int64_t c = bid * kWavesPerBlock + threadIdx.x / kWave;
  if (c >= p.n_chunks) return;
  const int64_t li = p.chunk_long[c];
  const int64_t row = p.long_rows[li];
  const int64_t c0 = p.long_chunk_ptr[li];
  const int64_t nc = p.long_chunk_ptr[li + 1] - c0;
  const int64_t rs = (int64_t)p.ptr[row], re = (int64_t)p.ptr[row + 1];
  const int64_t s = rs + (c - c0) * p.chunk_nnz;
  const int64_t e = s + p.chunk_nnz < re ? s + p.chunk_nnz : re;
  CA acc;
  acc.init();
  accumulate_span<kWave, U>(acc, p.idx, p.val, p.x, s, e, lane);
  group_reduce<kWave>(acc);
  if (nc == 1) {
    if (lane == 0) write_out(p.y, row, acc.value(), p.alpha, p.beta);
    return;
  }
  CA* parts = reinterpret_cast<CA*>(p.partials);
  int last = 0;
  if (lane == 0) {
    store_partial(parts + c, acc);
    const int old = __hip_atomic_fetch_add(p.counters + li, 1, __ATOMIC_ACQ_REL,
                                           __HIP_MEMORY_SCOPE_AGENT);
    last = (old == (int)nc - 1);
  }
  last = __builtin_amdgcn_readfirstlane(last);
  if (!last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  CA tot;
  tot.init();
  for (int64_t k = lane; k < nc; k += kWave) tot.merge(load_partial(parts + c0 + k));
  group_reduce<kWave>(tot);
  if (lane == 0) {
    write_out(p.y, row, tot.value(), p.alpha, p.beta);
    __hip_atomic_store(p.counters + li, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <typename V, typename A, typename I, typename P, int S, int U>
__global__ __launch_bounds__(kBlock) void spmv_csr_kernel(SpmvParams<V, I, P> p) {
  const int64_t bid = blockIdx.x;
  if (bid < p.n_chunk_blocks) {
    spmv_chunk_role<V, A, I, P, U>(p, bid);
    return;
  }
  int64_t sb = bid - p.n_chunk_blocks;
  if (p.xcd_remap) sb = xcd_remap_block(sb, (int64_t)gridDim.x - p.n_chunk_blocks);
  const int64_t row = sb * (kBlock / S) + threadIdx.x / S;
  if (row >= p.n_rows) return;  // whole S-group leaves together
  const int g = threadIdx.x % S;
  const int64_t rs = (int64_t)p.ptr[row], re = (int64_t)p.ptr[row + 1];
  const bool mine = (re - rs) <= p.short_max;
  A acc;
  acc.init();
  if (mine) accumulate_span<S, U>(acc, p.idx, p.val, p.x, rs, re, g);
  group_reduce<S>(acc);
  if (mine && g == 0) write_out(p.y, row, acc.value(), p.alpha, p.beta);
}

// Hot instantiations (fp32 / fp64 values, int32 columns) get every group
// width and unroll depth; the other dtype / index combinations are built for
// group widths {1, 4, 16, 64} and unroll 4 to bound the library size.
template <typename V, typename I>
constexpr bool kHotSpmv = (std::is_same<V, float>::value || std::is_same<V, double>::value) &&
                          std::is_same<I, int32_t>::value;

template <typename V, typename A, typename I, typename P, int S>
int launch_spmv_s(const fs_csr* M, const void* x, void* y, double alpha, double beta,
                  const fs_plan* pl, hipStream_t st) {
  SpmvParams<V, I, P> p;
  p.ptr = static_cast<const P*>(M->row_ptr);
  p.idx = static_cast<const I*>(M->col_idx);
  p.val = static_cast<const V*>(M->values);
  p.x = static_cast<const V*>(x);
  p.y = static_cast<V*>(y);
  p.n_rows = M->n_rows;
  p.alpha = alpha;
  p.beta = beta;
  p.short_max = pl->short_max;
  p.n_chunks = pl->n_chunks;
  p.n_chunk_blocks = (pl->n_chunks + kWavesPerBlock - 1) / kWavesPerBlock;
  p.chunk_nnz = pl->chunk_nnz;
  p.long_rows = pl->long_rows;
  p.long_chunk_ptr = pl->long_chunk_ptr;
  p.chunk_long = pl->chunk_long;
  p.counters = pl->counters;
  p.partials = pl->partials;
  p.xcd_remap = pl->xcd_remap;
  const int64_t rows_per_block = kBlock / S;
  const int64_t short_blocks = (M->n_rows + rows_per_block - 1) / rows_per_block;
  const int64_t grid = p.n_chunk_blocks + short_blocks;
  if (grid == 0) return FS_OK;
  if (grid > 0x7fffffffLL)
    return fail(FS_ERR_UNSUPPORTED, "fs_spmv_csr: grid of %lld blocks too large", (long long)grid);
  const dim3 g((unsigned)grid), b(kBlock);
  if constexpr (kHotSpmv<V, I>) {
    switch (pl->unroll) {
      case 2: spmv_csr_kernel<V, A, I, P, S, 2><<<g, b, 0, st>>>(p); break;
      case 8: spmv_csr_kernel<V, A, I, P, S, 8><<<g, b, 0, st>>>(p); break;
      default: spmv_csr_kernel<V, A, I, P, S, 4><<<g, b, 0, st>>>(p); break;
    }
  } else {
    spmv_csr_kernel<V, A, I, P, S, 4><<<g, b, 0, st>>>(p);
  }
  return check_launch("fs_spmv_csr");
}

template <typename V, typename A, typename I, typename P>
int launch_spmv(const fs_csr* M, const void* x, void* y, double alpha, double beta,
                const fs_plan* pl, hipStream_t st) {
  const int s = pl->group_lanes;
  if constexpr (kHotSpmv<V, I>) {
    switch (s) {
      case 1: return launch_spmv_s<V, A, I, P, 1>(M, x, y, alpha, beta, pl, st);
      case 2: return launch_spmv_s<V, A, I, P, 2>(M, x, y, alpha, beta, pl, st);
      case 4: return launch_spmv_s<V, A, I, P, 4>(M, x, y, alpha, beta, pl, st);
      case 8: return launch_spmv_s<V, A, I, P, 8>(M, x, y, alpha, beta, pl, st);
      case 16: return launch_spmv_s<V, A, I, P, 16>(M, x, y, alpha, beta, pl, st);
      case 32: return launch_spmv_s<V, A, I, P, 32>(M, x, y, alpha, beta, pl, st);
      case 64: return launch_spmv_s<V, A, I, P, 64>(M, x, y, alpha, beta, pl, st);
      default: break;
    }
  } else {
    if (s == 1 || s == 2) return launch_spmv_s<V, A, I, P, 1>(M, x, y, alpha, beta, pl, st);
    if (s == 4 || s == 8) return launch_spmv_s<V, A, I, P, 4>(M, x, y, alpha, beta, pl, st);
    if (s == 16 || s == 32) return launch_spmv_s<V, A, I, P, 16>(M, x, y, alpha, beta, pl, st);
    if (s == 64) return launch_spmv_s<V, A, I, P, 64>(M, x, y, alpha, beta, pl, st);
  }
  return fail(FS_ERR_VALUE, "fs_spmv_csr: group_lanes must be 1,2,4,8,16,32 or 64, got %d", s);
}

template <typename V, typename A>
int launch_spmv_idx(const fs_csr* M, const void* x, void* y, double alpha, double beta,
                    const fs_plan* pl, hipStream_t st) {
  if (M->index_type == FS_I32 && M->ptr_type == FS_I32) return launch_spmv<V, A, int32_t, int32_t>(M, x, y, alpha, beta, pl, st);
  if (M->index_type == FS_I32 && M->ptr_type == FS_I64) return launch_spmv<V, A, int32_t, int64_t>(M, x, y, alpha, beta, pl, st);
  if (M->index_type == FS_I64 && M->ptr_type == FS_I32) return launch_spmv<V, A, int64_t, int32_t>(M, x, y, alpha, beta, pl, st);
  return launch_spmv<V, A, int64_t, int64_t>(M, x, y, alpha, beta, pl, st);
}

}  // namespace fs

using namespace fs;

extern "C" int fs_spmv_csr(const fs_csr* A, const void* x, void* y, double alpha, double beta,
                           const fs_plan* plan, void* stream) {
  int rc = check_csr(A, "fs_spmv_csr");
  if (rc) return rc;
  rc = check_plan(plan, A, 1, "fs_spmv_csr");
  if (rc) return rc;
  if (A->n_rows == 0) return FS_OK;
  if (!y) return fail(FS_ERR_VALUE, "fs_spmv_csr: y is null");
  if (A->n_cols > 0 && !x) return fail(FS_ERR_VALUE, "fs_spmv_csr: x is null");
  hipStream_t st = as_stream(stream);
  const bool wide = plan->accumulate == FS_ACC_WIDE;
  switch (A->value_type) {
    case FS_F16: return launch_spmv_idx<f16_t, Plain<float>>(A, x, y, alpha, beta, plan, st);
    case FS_BF16: return launch_spmv_idx<bf16_t, Plain<float>>(A, x, y, alpha, beta, plan, st);
    case FS_F32:
      return wide ? launch_spmv_idx<float, Plain<double>>(A, x, y, alpha, beta, plan, st)
                  : launch_spmv_idx<float, Plain<float>>(A, x, y, alpha, beta, plan, st);
    case FS_F64: return launch_spmv_idx<double, Plain<double>>(A, x, y, alpha, beta, plan, st);
    case FS_C64: return launch_spmv_idx<c64_t, Plain<cplx<float>>>(A, x, y, alpha, beta, plan, st);
    case FS_C128: return launch_spmv_idx<c128_t, Plain<cplx<double>>>(A, x, y, alpha, beta, plan, st);
    case FS_C32: return launch_spmv_idx<c32_t, Plain<cplx<float>>>(A, x, y, alpha, beta, plan, st);
    default: return fail(FS_ERR_TYPE, "fs_spmv_csr: unsupported value_type %d", A->value_type);
  }
}
