// spmv_kernel.hpp -- the row-split CSR SpMV kernel, templated on how matrix
// values are stored (a "value source"), shared by the plain-storage entry point
// (spmv.hip) and the split-precision storage modes (precision.hip).
//
// One launch, two block roles:
//   * chunk blocks (lowest block ids, so the long work starts first): each
//     wavefront reduces one chunk of a listed long row; a row split over
//     several chunks is finished by whichever wavefront arrives last, which
//     sums the row's partials in chunk order -- the result does not depend on
//     arrival order.
//   * short blocks: groups of S lanes (S = 1..64) take consecutive rows in
//     natural order; each lane keeps U entries in flight (coalesced col/val
//     streams, U independent x gathers), then the group reduces with DPP /
//     swizzle exchanges.  Rows longer than short_max are skipped (chunked).
#pragma once

#include <type_traits>

#include "fs_host.hpp"

namespace fs {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;

// Value source of plainly stored values of type V.
template <typename V>
struct DirectVals {
  const V* v;
  template <typename C> __device__ __forceinline__ C get(int64_t j) const { return load_as<C>(v, j); }
  static constexpr bool kRowScale = false;
  __device__ __forceinline__ float row_scale(int64_t) const { return 1.f; }
};

template <typename VS, typename V, typename I, typename P>
struct SpmvParams {
  const P* ptr;
  const I* idx;
  VS vals;
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

// Accumulate entries [s, e) of one row with STRIDE lanes, lane offset g,
// U entries in flight per lane.
template <int STRIDE, int U, typename A, typename VS, typename V, typename I>
__device__ __forceinline__ void accumulate_span(A& acc, const I* __restrict__ idx, const VS& vals,
                                                const V* __restrict__ x, int64_t s, int64_t e,
                                                int g) {
  using C = typename A::compute_t;
  int64_t j = s + g;
  for (; j + (int64_t)(U - 1) * STRIDE < e; j += (int64_t)U * STRIDE) {
    int64_t col[U];
    C a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      col[u] = ldg_index(idx, j + u * STRIDE);
      a[u] = vals.template get<C>(j + u * STRIDE);
    }
    C xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xv[u] = load_as<C>(x, col[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) acc.add_prod(a[u], xv[u]);
  }
  for (; j < e; j += STRIDE) {
    const int64_t c = ldg_index(idx, j);
    acc.add_prod(vals.template get<C>(j), load_as<C>(x, c));
  }
}

template <typename VS, typename V, typename T>
__device__ __forceinline__ void write_row(const VS& vals, V* y, int64_t row, T v, double alpha,
                                          double beta) {
  if constexpr (VS::kRowScale) alpha *= (double)vals.row_scale(row);
  write_out(y, row, v, alpha, beta);
}

template <typename A, int U, typename VS, typename V, typename I, typename P>
__device__ __forceinline__ void spmv_chunk_role(const SpmvParams<VS, V, I, P>& p, int64_t bid) {
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
  accumulate_span<kWave, U>(acc, p.idx, p.vals, p.x, s, e, lane);
  group_reduce<kWave>(acc);
  if (nc == 1) {
    if (lane == 0) write_row(p.vals, p.y, row, acc.value(), p.alpha, p.beta);
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
    write_row(p.vals, p.y, row, tot.value(), p.alpha, p.beta);
    __hip_atomic_store(p.counters + li, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <typename A, int S, int U, typename VS, typename V, typename I, typename P>
__device__ __forceinline__ void spmv_short_role(const SpmvParams<VS, V, I, P>& p, int64_t sb) {
  const int64_t row = sb * (kBlock / S) + threadIdx.x / S;
  if (row >= p.n_rows) return;  // whole S-group leaves together
  const int g = threadIdx.x % S;
  const int64_t rs = (int64_t)p.ptr[row], re = (int64_t)p.ptr[row + 1];
  const bool mine = (re - rs) <= p.short_max;
  A acc;
  acc.init();
  if (mine) accumulate_span<S, U>(acc, p.idx, p.vals, p.x, rs, re, g);
  group_reduce<S>(acc);
  if (mine && g == 0) write_row(p.vals, p.y, row, acc.value(), p.alpha, p.beta);
}

template <typename A, int S, int U, typename VS, typename V, typename I, typename P>
__global__ __launch_bounds__(kBlock) void spmv_csr_kernel(SpmvParams<VS, V, I, P> p) {
  const int64_t bid = blockIdx.x;
  if (bid < p.n_chunk_blocks) {
    spmv_chunk_role<A, U>(p, bid);
    return;
  }
  int64_t sb = bid - p.n_chunk_blocks;
  if (p.xcd_remap) sb = xcd_remap_block(sb, (int64_t)gridDim.x - p.n_chunk_blocks);
  spmv_short_role<A, S, U>(p, sb);
}

// Fills the schedule fields of SpmvParams from a plan; returns the grid size.
template <typename VS, typename V, typename I, typename P>
int64_t spmv_fill_schedule(SpmvParams<VS, V, I, P>& p, const fs_csr* M, const fs_plan* pl,
                           int S) {
  p.n_rows = M->n_rows;
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
  return p.n_chunk_blocks + (M->n_rows + rows_per_block - 1) / rows_per_block;
}

}  // namespace fs
