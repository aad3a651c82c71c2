// spmv_kernel.hpp -- the row-split CSR SpMV kernel, templated on how matrix
// values are stored (a "value source"), shared by the plain-storage entry point
// (spmv.hip) and the split-precision storage modes (precision.hip).
//
// One launch, two block roles:
//   * chunk blocks (lowest block ids, so the long work starts first): each
//     wavefront reduces one chunk of a listed long row; a row that fits one
//     chunk is written directly, otherwise the chunk's partial is stored and
//     spmv_fixup_kernel (a second, small launch) sums the row's partials in
//     chunk order -- deterministic, and no cross-workgroup synchronisation
//     inside a launch (agent-scope release / acquire on gfx950 writes back /
//     invalidates L2, which costs more than the extra launch).
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
  void* partials;
  int xcd_remap;
  int64_t tile_nnz;            // tile role: target entries per tile (0: row-group role)
  int64_t n_tiles;
  const int64_t* tile_start;   // [n_tiles + 1] first row of each tile
};

// Accumulator used for chunks: fp32-native chunked rows get compensation.
template <typename A> struct ChunkAcc { using T = A; };
template <> struct ChunkAcc<Plain<float>> { using T = Comp; };

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
  if (lane != 0) return;
  if (nc == 1)
    write_row(p.vals, p.y, row, acc.value(), p.alpha, p.beta);
  else
    reinterpret_cast<CA*>(p.partials)[c] = acc;  // summed by spmv_fixup_kernel
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

// ------------------------------------------------------------- tile role
//
// A block takes a tile of consecutive short rows holding about tile_nnz
// entries (at most 2 * tile_nnz, at most kTileRows rows; a tile that is a
// single long row is skipped -- the chunk role owns it).  The tile's row
// pointers are staged in LDS; every thread then streams U entries at a time
// of the tile's contiguous entry range (coalesced col / value loads, U
// independent x gathers in flight), and writes the products to LDS.  Rows are
// then summed from LDS in entry order: short rows by one thread each, rows
// over kThreadRowMax entries by a whole wavefront.  Row sums are carried in
// fp64 (complex: two fp64), so a row's value does not depend on the tile
// layout beyond the product rounding of native fp32.

constexpr int kTileRows = 512;
constexpr int kThreadRowMax = 32;

template <typename C> struct TileSum { using T = double; };
template <> struct TileSum<cplx<float>> { using T = cplx<double>; };
template <> struct TileSum<cplx<double>> { using T = cplx<double>; };

__device__ __forceinline__ double widen(float v) { return (double)v; }
__device__ __forceinline__ double widen(double v) { return v; }
template <typename T> __device__ __forceinline__ cplx<double> widen(cplx<T> v) {
  return {(double)v.re, (double)v.im};
}
__device__ __forceinline__ double mulc(double a, double b) { return a * b; }
template <typename T> __device__ __forceinline__ cplx<T> mulc(cplx<T> a, cplx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ float mulc(float a, float b) { return a * b; }

// Bytes of dynamic LDS the tile role needs for a product type S.
template <typename S> __host__ __device__ inline size_t tile_lds_bytes(int64_t tile_nnz) {
  return sizeof(int64_t) * (kTileRows + 2) + sizeof(S) * 2 * (size_t)tile_nnz;
}

template <typename A, int U, typename VS, typename V, typename I, typename P>
__device__ __forceinline__ void spmv_tile_role(const SpmvParams<VS, V, I, P>& p, int64_t t,
                                               char* smem) {
  using C = typename A::compute_t;
  using S = typename TileSum<C>::T;
  int64_t* rp = reinterpret_cast<int64_t*>(smem);
  S* prod = reinterpret_cast<S*>(smem + sizeof(int64_t) * (kTileRows + 2));
  const int64_t r0 = p.tile_start[t];
  int64_t r1 = p.tile_start[t + 1];
  if (r1 - r0 > kTileRows) r1 = r0 + kTileRows;  // plans never exceed it; stay in LDS
  const int nr = (int)(r1 - r0);
  for (int i = threadIdx.x; i <= nr; i += kBlock) rp[i] = (int64_t)p.ptr[r0 + i];
  __syncthreads();
  const int64_t e0 = rp[0], e1 = rp[nr];
  if (nr == 1 && e1 - e0 > p.short_max) return;  // a long row: the chunk role's
  if (e1 - e0 > 2 * p.tile_nnz) {
    // a tile the plan builder would not make (too many entries for LDS):
    // reduce its rows one wavefront each straight from global memory
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    for (int i = w; i < nr; i += kWavesPerBlock) {
      A acc;
      acc.init();
      accumulate_span<kWave, U>(acc, p.idx, p.vals, p.x, rp[i], rp[i + 1], lane);
      group_reduce<kWave>(acc);
      if (lane == 0) write_row(p.vals, p.y, r0 + i, acc.value(), p.alpha, p.beta);
    }
    return;
  }
  // products of the tile's entries, U per thread in flight
  for (int64_t base = e0 + threadIdx.x; base < e1; base += (int64_t)kBlock * U) {
    int64_t col[U];
    C a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t j = base + (int64_t)u * kBlock;
      col[u] = j < e1 ? ldg_index(p.idx, j) : 0;
      a[u] = j < e1 ? p.vals.template get<C>(j) : zero<C>();
    }
    C xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t j = base + (int64_t)u * kBlock;
      xv[u] = j < e1 ? load_as<C>(p.x, col[u]) : zero<C>();
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t j = base + (int64_t)u * kBlock;
      if (j < e1) prod[j - e0] = widen(mulc(a[u], xv[u]));
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nr; i += kBlock) {
    const int64_t s = rp[i] - e0, e = rp[i + 1] - e0;
    if (e - s > kThreadRowMax) continue;
    S acc = zero<S>();
    for (int64_t j = s; j < e; ++j) acc = add(acc, prod[j]);
    write_row(p.vals, p.y, r0 + i, acc, p.alpha, p.beta);
  }
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  for (int i = w; i < nr; i += kWavesPerBlock) {
    const int64_t s = rp[i] - e0, e = rp[i + 1] - e0;
    if (e - s <= kThreadRowMax) continue;
    Plain<S> acc;
    acc.init();
    for (int64_t j = s + lane; j < e; j += kWave) acc.add_value(prod[j]);
    group_reduce<kWave>(acc);
    if (lane == 0) write_row(p.vals, p.y, r0 + i, acc.value(), p.alpha, p.beta);
  }
}

template <typename A, int U, typename VS, typename V, typename I, typename P>
__global__ __launch_bounds__(kBlock) void spmv_tile_kernel(SpmvParams<VS, V, I, P> p) {
  extern __shared__ __align__(16) char smem[];
  const int64_t bid = blockIdx.x;
  if (bid < p.n_chunk_blocks) {
    spmv_chunk_role<A, U>(p, bid);
    return;
  }
  int64_t t = bid - p.n_chunk_blocks;
  if (p.xcd_remap) t = xcd_remap_block(t, (int64_t)gridDim.x - p.n_chunk_blocks);
  spmv_tile_role<A, U>(p, t, smem);
}

// Second pass for rows cut into several chunks: one wavefront per long row
// sums the row's chunk partials in a fixed order (lane k takes chunks k,
// k + 64, ...; then the wave tree) and writes the row.  Runs after the main
// kernel on the same stream, so every partial is visible without any
// cross-workgroup synchronisation inside a launch.
template <typename A, typename VS, typename V, typename I, typename P>
__global__ __launch_bounds__(kBlock) void spmv_fixup_kernel(SpmvParams<VS, V, I, P> p,
                                                            int64_t n_long) {
  using CA = typename ChunkAcc<A>::T;
  const int lane = threadIdx.x % kWave;
  const int64_t li = (int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (li >= n_long) return;
  const int64_t c0 = p.long_chunk_ptr[li];
  const int64_t nc = p.long_chunk_ptr[li + 1] - c0;
  if (nc <= 1) return;  // written by the main kernel
  const CA* parts = reinterpret_cast<const CA*>(p.partials);
  CA tot;
  tot.init();
  for (int64_t k = lane; k < nc; k += kWave) tot.merge(parts[c0 + k]);
  group_reduce<kWave>(tot);
  if (lane == 0) write_row(p.vals, p.y, p.long_rows[li], tot.value(), p.alpha, p.beta);
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
  p.partials = pl->partials;
  p.xcd_remap = pl->xcd_remap;
  p.tile_nnz = pl->tile_nnz;
  p.n_tiles = pl->n_tiles;
  p.tile_start = pl->tile_start;
  if (pl->tile_nnz > 0) return p.n_chunk_blocks + pl->n_tiles;
  const int64_t rows_per_block = kBlock / S;
  return p.n_chunk_blocks + (M->n_rows + rows_per_block - 1) / rows_per_block;
}

// Launches the tile-role kernel (chunk + tile blocks) with its dynamic LDS.
template <typename A, int U, typename VS, typename V, typename I, typename P>
void spmv_launch_tiles(const SpmvParams<VS, V, I, P>& p, int64_t grid, hipStream_t st) {
  using S = typename TileSum<typename A::compute_t>::T;
  spmv_tile_kernel<A, U><<<dim3((unsigned)grid), dim3(kBlock), tile_lds_bytes<S>(p.tile_nnz), st>>>(p);
}

// Launches the fix-up pass when some long row spans several chunks.
template <typename A, typename VS, typename V, typename I, typename P>
int spmv_launch_fixup(const SpmvParams<VS, V, I, P>& p, const fs_plan* pl, hipStream_t st,
                      const char* who) {
  if (pl->n_chunks <= pl->n_long) return FS_OK;  // every long row fits one chunk
  const int64_t blocks = (pl->n_long + kWavesPerBlock - 1) / kWavesPerBlock;
  spmv_fixup_kernel<A><<<dim3((unsigned)blocks), dim3(kBlock), 0, st>>>(p, pl->n_long);
  return check_launch(who);
}

}  // namespace fs
