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

typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

// Four consecutive elements p[j .. j+3] with one 8-byte or one to four 16-byte
// loads; needs j % 4 == 0 and p aligned to min(16, 4 * sizeof(T)) bytes.
// Every caller streams an array once (entry-tile values, columns, x
// permutations, split-precision planes, gathered x), so the loads are
// non-temporal: they do not push the lines that are re-read (x gathers, the
// shared ends of gathered-x runs) out of L2.  On the 4M-row power-law SpMV
// this took the entry-tile pass from 121 to 105 us, and the 4M-row banded
// SpMV from 89.7 to 83.7 us.
template <typename T>
__device__ __forceinline__ void load4(const T* __restrict__ p, int64_t j, T (&o)[4]) {
  static_assert(sizeof(T) == 2 || sizeof(T) % 4 == 0, "element size");
  if constexpr (sizeof(T) == 2) {
    const u32x2_t v = __builtin_nontemporal_load(reinterpret_cast<const u32x2_t*>(p + j));
    __builtin_memcpy(o, &v, 8);
  } else {
    constexpr int N = (int)sizeof(T) * 4 / 16;
    const u32x4_t* q = reinterpret_cast<const u32x4_t*>(p + j);
    u32x4_t v[N];
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = __builtin_nontemporal_load(q + k);
    __builtin_memcpy(o, v, sizeof(T) * 4);
  }
}

// Value source of plainly stored values of type V.
template <typename V>
struct DirectVals {
  const V* v;
  template <typename C> __device__ __forceinline__ C get(int64_t j) const { return load_as<C>(v, j); }
  // vector form for the entry-tile role (j % 4 == 0, v aligned)
  static constexpr bool kVec = true;
  template <typename C> __device__ __forceinline__ void get4(int64_t j, C (&o)[4]) const {
    V t[4];
    load4(v, j, t);
#pragma unroll
    for (int k = 0; k < 4; ++k) o[k] = load_as<C>(t, k);
  }
  static constexpr bool kRowScale = false;
  __device__ __forceinline__ float row_scale(int64_t) const { return 1.f; }
  // whether get4's loads are aligned (host side)
  bool vec_aligned() const {
    const size_t va = 4 * sizeof(V) >= 16 ? 16 : 4 * sizeof(V);
    return ((uintptr_t)v % va) == 0;
  }
  // raw stored value and its conversion, for loads that must stay
  // unconverted until a whole batch is issued (accumulate_span)
  static constexpr bool kRaw = std::is_arithmetic<V>::value;
  using raw_t = V;
  __device__ __forceinline__ V raw(int64_t j) const { return v[j]; }
  template <typename C> __device__ __forceinline__ C conv(V r) const { return (C)r; }
};

template <typename VS, typename = void> struct HasRaw { static constexpr bool value = false; };
template <typename VS> struct HasRaw<VS, std::void_t<decltype(VS::kRaw)>> {
  static constexpr bool value = VS::kRaw;
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
  const int64_t* tile_entry;   // entry tiles: [n_tiles + 1] first entry of each tile
  int64_t nnz;
  int n_groups;                // entry tiles: block groups (0: block b = tile b)
  const int64_t* group_tiles;  // [n_groups + 1]
  void* row_sums;              // entry tiles: raw row sums instead of y (or null)
  // gathered x stored window-major: the run table of the tile-order positions
  const int32_t* xg_tile_run;  // [n_tiles + 1] first run of each tile
  const int32_t* xg_batch_run;
  const uint64_t* xg_batch_mask;
  const int32_t* xg_run_off;   // null: x is tile-major (one stream per tile)
  int64_t xg_n_runs;
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
  if (j < e) {
    // the last, partial batch (a row shorter than U * STRIDE is only this):
    // all of its loads in flight at once, entries past e masked; the
    // products are added in the same order as one at a time.  Each masked
    // load's value is kept raw and opaque (empty asm) until the batch is
    // issued: a conversion of it (index widening, fp32 -> fp64) would be
    // placed inside the load's branch together with a wait for the load,
    // serialising the batch.
    if constexpr (HasRaw<VS>::value && std::is_arithmetic<V>::value) {
      using R = typename VS::raw_t;
      I ci[U];
      R ar[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool ok = j + u * STRIDE < e;
        ci[u] = ok ? idx[j + u * STRIDE] : I(0);
        ar[u] = ok ? vals.raw(j + u * STRIDE) : R(0);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) asm volatile("" : "+v"(ci[u]));
      V xr[U];
#pragma unroll
      for (int u = 0; u < U; ++u) xr[u] = j + u * STRIDE < e ? x[(int64_t)ci[u]] : V(0);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        asm volatile("" : "+v"(ar[u]));
        asm volatile("" : "+v"(xr[u]));
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (j + u * STRIDE < e) acc.add_prod(vals.template conv<C>(ar[u]), (C)xr[u]);
      return;
    }
    I ci[U];
    C a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool ok = j + u * STRIDE < e;
      ci[u] = ok ? idx[j + u * STRIDE] : I(0);
      a[u] = ok ? vals.template get<C>(j + u * STRIDE) : zero<C>();
    }
#pragma unroll
    for (int u = 0; u < U; ++u) asm volatile("" : "+v"(ci[u]));
    C xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      xv[u] = j + u * STRIDE < e ? load_as<C>(x, (int64_t)ci[u]) : zero<C>();
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (j + u * STRIDE < e) acc.add_prod(a[u], xv[u]);
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

// Row-sum type of an accumulator's tile roles: fp64 (two fp64 for complex)
// for Plain accumulators, a compensated fp32 pair for Comp (the split-
// precision modes, which accumulate in fp32 only).
template <typename A> struct TileSumOf { using T = typename TileSum<typename A::compute_t>::T; };
template <> struct TileSumOf<Comp> { using T = f32c; };

// A product or partial as a row-sum term.
template <typename S, typename C> __device__ __forceinline__ S to_sum(C v) { return widen(v); }
template <> __device__ __forceinline__ f32c to_sum<f32c, float>(float v) { return {v, 0.f}; }
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
  using S = typename TileSumOf<A>::T;
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
    if constexpr (HasRaw<VS>::value && std::is_arithmetic<V>::value) {
      // raw loads kept opaque until all are issued, then converted (a
      // conversion inside a masked load's branch brings a wait with it)
      using R = typename VS::raw_t;
      I ci[U];
      R ar[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t j = base + (int64_t)u * kBlock;
        ci[u] = j < e1 ? p.idx[j] : I(0);
        ar[u] = j < e1 ? p.vals.raw(j) : R(0);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) asm volatile("" : "+v"(ci[u]));
      V xr[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t j = base + (int64_t)u * kBlock;
        xr[u] = j < e1 ? p.x[(int64_t)ci[u]] : V(0);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        asm volatile("" : "+v"(ar[u]));
        asm volatile("" : "+v"(xr[u]));
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t j = base + (int64_t)u * kBlock;
        if (j < e1) prod[j - e0] = to_sum<S>(mulc(p.vals.template conv<C>(ar[u]), (C)xr[u]));
      }
      continue;
    }
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
      if (j < e1) prod[j - e0] = to_sum<S>(mulc(a[u], xv[u]));
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nr; i += kBlock) {
    const int64_t s = rp[i] - e0, e = rp[i + 1] - e0;
    if (e - s > kThreadRowMax) continue;
    S acc = zero<S>();
    for (int64_t j = s; j < e; ++j) acc = add(acc, prod[j]);
    write_row(p.vals, p.y, r0 + i, final_value(acc), p.alpha, p.beta);
  }
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  for (int i = w; i < nr; i += kWavesPerBlock) {
    const int64_t s = rp[i] - e0, e = rp[i + 1] - e0;
    if (e - s <= kThreadRowMax) continue;
    Plain<S> acc;
    acc.init();
    for (int64_t j = s + lane; j < e; j += kWave) acc.add_value(prod[j]);
    group_reduce<kWave>(acc);
    if (lane == 0) write_row(p.vals, p.y, r0 + i, final_value(acc.value()), p.alpha, p.beta);
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
  p.tile_entry = pl->tile_entry;
  p.nnz = M->nnz;
  p.n_groups = pl->n_groups;
  p.group_tiles = pl->group_tiles;
  p.row_sums = pl->row_sums;
  p.xg_tile_run = pl->xg_tile_run;
  p.xg_batch_run = pl->xg_batch_run;
  p.xg_batch_mask = pl->xg_batch_mask;
  p.xg_run_off = pl->xg_run_off;
  p.xg_n_runs = pl->xg_n_runs;
  if (pl->tile_nnz > 0 && pl->tile_mode == FS_TILE_ENTRIES) {
    if (pl->n_groups <= 0) return pl->n_tiles;
    return pl->group_blocks * pl->n_groups;
  }
  if (pl->tile_nnz > 0) return p.n_chunk_blocks + pl->n_tiles;
  const int64_t rows_per_block = kBlock / S;
  return p.n_chunk_blocks + (M->n_rows + rows_per_block - 1) / rows_per_block;
}

// Launches the tile-role kernel (chunk + tile blocks) with its dynamic LDS.
template <typename A, int U, typename VS, typename V, typename I, typename P>
void spmv_launch_tiles(const SpmvParams<VS, V, I, P>& p, int64_t grid, hipStream_t st) {
  using S = typename TileSumOf<A>::T;
  spmv_tile_kernel<A, U><<<dim3((unsigned)grid), dim3(kBlock), tile_lds_bytes<S>(p.tile_nnz), st>>>(p);
}

// ------------------------------------------------------------ entry tiles
//
// Tile t holds the entries [tile_entry[t], tile_entry[t+1]) -- at most T =
// tile_nnz = E * kBlock of them -- and owns the rows tile_start[t] ..
// tile_start[t+1] - 1, which start inside its entry range (the last tile also
// owns the rows that start at nnz).  The plan cuts the entries every T and
// cuts again every kSegRows owned rows, so a run of empty rows is spread over
// many blocks.  The entry range is known from two scalar loads, so every
// thread issues all E of its col / value loads (16-byte vector loads when VEC
// and the tile is full and aligned) before the first x gather, and the first
// batch of row pointers is loaded beside them.  Products
// go to LDS in entry order; every owned row is summed there in fp64 (complex:
// two fp64) in entry order -- one thread per row up to kThreadRowMax entries,
// one wavefront above.  Two pieces can run past a tile: the head (entries of
// the last row of an earlier tile) and the tile's own last row; their sums go
// to carries[2t] and carries[2t+1], and spmv_entry_fixup_kernel adds a
// straddling row's carries in tile order.  Rows of any length take the same
// path; results depend on T only through where a long row is cut.

// Row boundaries staged in LDS per pass.  768 (not 1024) keeps a 4096-entry
// fp32 tile's LDS (16 KB of products + these slots) within 20 KB, so 8
// blocks -- the 32-wave limit -- fit a CU instead of 7: on the 4M-row
// power-law SpMV the gathered-x pass went from 104 to 100 us.
constexpr int kSegRows = 768;
// Rows longer than kThreadRowMax inside one tile: at most tile_nnz / 33 (124
// for the largest tile, 4096 entries), listed in LDS for the wavefronts.
constexpr int kLongRowsMax = 128;
// int slots after the products: [0, kSegRows] row boundaries, then the
// straddle flag, the long-row count, and from kLongList the long-row list
// (a multiple of 4 slots, so what follows stays 16-byte aligned).
constexpr int kStraddleSlot = kSegRows + 1;
constexpr int kLongCountSlot = kSegRows + 2;
constexpr int kLongList = kSegRows + 4;
constexpr int kBndSlots = kLongList + kLongRowsMax;

template <typename C> __host__ __device__ inline size_t entry_lds_bytes(int64_t tile_nnz) {
  using S = typename TileSum<C>::T;
  return sizeof(C) * (size_t)tile_nnz + sizeof(int32_t) * kBndSlots + sizeof(S) * kWavesPerBlock;
}

// Gathered-x form (fs_plan.x_perm): the tile's gathered x values (type V)
// are staged where the products go (they are read into registers, then a
// barrier, before any product is written) when V is no wider than the
// product type C; otherwise they follow the entry role's LDS, 16-byte aligned.
template <typename C, typename V> constexpr bool kXgAliased = sizeof(V) <= sizeof(C);
template <typename C, typename V> __host__ __device__ inline size_t entry_xg_offset(int64_t tile_nnz) {
  if constexpr (kXgAliased<C, V>) return 0;
  return (entry_lds_bytes<C>(tile_nnz) + 15) & ~(size_t)15;
}
template <typename C, typename V> __host__ __device__ inline size_t entry_xg_lds_bytes(int64_t tile_nnz) {
  if constexpr (kXgAliased<C, V>) return entry_lds_bytes<C>(tile_nnz);
  return entry_xg_offset<C, V>(tile_nnz) + sizeof(V) * (size_t)tile_nnz;
}
// Window-major gathered x also stages the tile's run table (XgwLds: batch
// records of T / 64 + 2 batches, T run offsets), which is dead once the
// tile's x values are loaded: the run offsets go where the products go
// (4 T <= sizeof(C) T bytes) and the batch records in the row-boundary
// slots, neither of which is written before the products; a barrier after
// the last read of the table lets the staged x values overwrite it.  No LDS
// beyond the tile-major form's.
template <typename C, typename V> __host__ __device__ inline size_t entry_xgw_lds_bytes(int64_t tile_nnz) {
  return entry_xg_lds_bytes<C, V>(tile_nnz);
}
static_assert(sizeof(int32_t) * kBndSlots >= 12 * (4096 / 64 + 2),
              "the row-boundary slots hold the batch records of a 4096-entry tile");
// LDS of the entry kernel: XG 0 plain, 1 tile-major gathered x, 2 window-major.
template <typename C, typename V> __host__ __device__ inline size_t entry_kernel_lds(int xg,
                                                                                    int64_t tile_nnz) {
  return xg == 2 ? entry_xgw_lds_bytes<C, V>(tile_nnz)
                 : xg == 1 ? entry_xg_lds_bytes<C, V>(tile_nnz) : entry_lds_bytes<C>(tile_nnz);
}

// A finished row of the entry role: y = alpha * sum + beta * y, or, with
// row_sums set, the raw sum in the sum type (times the row scale of scaled
// storage) for a later pass to combine.
template <typename VS, typename V, typename I, typename P, typename S>
__device__ __forceinline__ void entry_write(const SpmvParams<VS, V, I, P>& p, int64_t row, S acc) {
  if (p.row_sums) {
    if constexpr (VS::kRowScale) acc = scale(acc, (double)p.vals.row_scale(row));
    static_cast<S*>(p.row_sums)[row] = acc;
  } else {
    write_row(p.vals, p.y, row, final_value(acc), p.alpha, p.beta);
  }
}

template <typename A, typename VS, typename V, typename I, typename P, typename S>
__device__ __forceinline__ void entry_finish(const SpmvParams<VS, V, I, P>& p, int64_t row,
                                             S acc, bool to_carry, S* carries, int64_t t) {
  if (to_carry) carries[2 * t + 1] = acc;
  else entry_write(p, row, acc);
}

// One entry tile held in registers between its loads and its LDS stage.
template <typename I, typename C, int E> struct EntryTile {
  static constexpr int kPv = kSegRows / kBlock + 1;
  I col[E];
  C a[E];
  int64_t pv[kPv];  // first batch of row pointers (i <= nb0)
  int64_t t, e0, r0, r1;
  int len, nb0;
  bool vec;  // loaded with the 4-entry vector layout
};

// Bounds of tile t and every load that does not depend on another load:
// the col / value entries (16-byte vector loads when the tile is full and
// aligned, else scalar loads masked by the tile length) and the first batch
// of row pointers.
template <int E, bool VEC, typename VS, typename V, typename I, typename P, typename C>
__device__ __forceinline__ void entry_fetch(const SpmvParams<VS, V, I, P>& p, int64_t t,
                                            EntryTile<I, C, E>& f) {
  constexpr int T = E * kBlock;
  const int tid = threadIdx.x;
  f.t = t;
  // entry range [e0, e0 + len): len <= T (clamped, so a damaged plan cannot
  // overrun the LDS products)
  f.e0 = p.tile_entry[t];
  int64_t e1 = p.tile_entry[t + 1];
  e1 = e1 > p.nnz ? p.nnz : e1;
  const int64_t span = e1 - f.e0;
  f.len = span <= 0 ? 0 : (span < T ? (int)span : T);
  // owned rows [r0, r1), clamped to [0, n_rows] so a damaged plan cannot
  // send a row-pointer read out of bounds
  int64_t r0 = p.tile_start[t], r1 = p.tile_start[t + 1];
  r0 = r0 < 0 ? 0 : (r0 > p.n_rows ? p.n_rows : r0);
  r1 = r1 < r0 ? r0 : (r1 > p.n_rows ? p.n_rows : r1);
  f.r0 = r0;
  f.r1 = r1;
  f.nb0 = (int)(r1 - r0 < kSegRows ? r1 - r0 : kSegRows);
  // Row pointers: a batch wholly past the owned rows is skipped by the whole
  // block (nb0 is uniform); inside a batch, lanes past nb0 re-read the last
  // owned pointer (r0 + nb0 <= r1 <= n_rows) instead of branching per lane --
  // the compiler waits for a load under a per-lane branch inside the branch,
  // which would serialise these loads with each other and with the entry
  // loads below.
  constexpr int kPv = EntryTile<I, C, E>::kPv;
  P pr[kPv];
#pragma unroll
  for (int k = 0; k < kPv; ++k) {
    const int i = k * kBlock + tid;
    pr[k] = k * kBlock <= f.nb0 ? p.ptr[r0 + (i <= f.nb0 ? i : f.nb0)] : P(0);
  }
  f.vec = VEC && f.len == T && (f.e0 & 3) == 0;
  if (f.vec) {
    if constexpr (VEC) {
#pragma unroll
      for (int g = 0; g < E / 4; ++g) {
        const int64_t j = f.e0 + (int64_t)(g * kBlock + tid) * 4;
        I ci[4];
        load4(p.idx, j, ci);
        C av[4];
        p.vals.template get4<C>(j, av);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          f.col[4 * g + k] = ci[k];
          f.a[4 * g + k] = av[k];
        }
      }
    }
  } else if (f.len == T) {
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int64_t j = f.e0 + u * kBlock + tid;
      f.col[u] = p.idx[j];
      f.a[u] = p.vals.template get<C>(j);
    }
  } else {
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int o = u * kBlock + tid;
      const int64_t j = f.e0 + o;
      f.col[u] = o < f.len ? p.idx[j] : I(0);
      f.a[u] = o < f.len ? p.vals.template get<C>(j) : zero<C>();
    }
  }
  // the row pointers are widened only now that the entry loads are issued;
  // the empty asm keeps the widening (and its wait) from moving up into the
  // row-pointer branches
#pragma unroll
  for (int k = 0; k < kPv; ++k) {
    asm volatile("" : "+v"(pr[k]));
    f.pv[k] = (int64_t)pr[k];
  }
}

// The x gathers of a fetched tile (masked past its length).
template <int E, typename VS, typename V, typename I, typename P, typename C>
__device__ __forceinline__ void entry_gather(const SpmvParams<VS, V, I, P>& p,
                                             const EntryTile<I, C, E>& f, C (&xv)[E]) {
  constexpr int T = E * kBlock;
  if (f.len == T) {
#pragma unroll
    for (int u = 0; u < E; ++u) xv[u] = load_as<C>(p.x, (int64_t)f.col[u]);
  } else {  // a partial tile is never in the vector layout
    const int tid = threadIdx.x;
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int o = u * kBlock + tid;
      xv[u] = o < f.len ? load_as<C>(p.x, (int64_t)f.col[u]) : zero<C>();
    }
  }
}

// Gathered-x form: the tile's x values, already gathered per entry, lie at
// x[e0 .. e0 + len) in the order of x_perm (f.col holds each entry's x_perm
// value).  They are staged in LDS with coalesced loads -- all issued before
// the first LDS store -- and then read through the permutation (clamped into
// the tile, so a damaged permutation cannot leave it).

template <int E, typename VS, typename V, typename I, typename P, typename C>
__device__ __forceinline__ void entry_stage_xg(const SpmvParams<VS, V, I, P>& p,
                                               const EntryTile<I, C, E>& f, V* xt) {
  const int tid = threadIdx.x;
  if (f.len == 0) return;  // uniform: an empty tile (e0 may equal nnz) reads nothing
  if (f.vec) {
#pragma unroll
    for (int g = 0; g < E / 4; ++g) {
      const int o = (g * kBlock + tid) * 4;
      V v[4];
      load4(p.x, f.e0 + o, v);
#pragma unroll
      for (int k = 0; k < 4; ++k) xt[o + k] = v[k];
    }
  } else {
    V v[E];
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int o = u * kBlock + tid;
      v[u] = p.x[f.e0 + (o < f.len ? o : 0)];
    }
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int o = u * kBlock + tid;
      if (o < f.len) xt[o] = v[u];
    }
  }
}

// Window-major gathered x (fs_plan.xg_run_off): tile position d = e0 + o
// lies in run r = batch_run[d / 64] + (run starts in the batch at 1 .. d % 64)
// and its value at x[d + run_off[r]].  The tile's batch records and the run
// offsets of its runs [tile_run[t], tile_run[t + 1]) -- at most T / 64 + 1
// batches and T runs -- are loaded into LDS with one round of loads (each
// thread loads a clamped index, so no load waits behind a per-lane branch;
// runs past the first kBlock, which take windows narrower than T / kBlock
// entries per tile, follow in a loop); after a barrier every position's value
// load is issued at once, and after a second one the values are stored.  The
// table's loads are issued before the tile's entry loads and both barriers
// order LDS only, so the entry loads stay in flight throughout.  Consecutive
// positions of a run are consecutive in x, so a wavefront's loads coalesce
// run by run.  Every index is clamped into its array.
template <int T> struct XgwLds {
  static constexpr int kBatches = T / kWave + 2;
  uint64_t* mask;
  int32_t* run;
  int32_t* off;
  __device__ __forceinline__ XgwLds(char* batches, char* offsets)
      : mask(reinterpret_cast<uint64_t*>(batches)),
        run(reinterpret_cast<int32_t*>(batches + sizeof(uint64_t) * kBatches)),
        off(reinterpret_cast<int32_t*>(offsets)) {}
};

// A barrier that orders LDS only: outstanding global loads stay in flight
// across it (__syncthreads waits for every outstanding load as well).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The tile's run-table loads, issued before the tile's entry loads so that
// waiting for them does not wait for those.
struct XgwFetch {
  int64_t e0, q0, b0;
  int len, nr, nb;
  int32_t br, ro;
  uint64_t bm;
};

template <int E, typename VS, typename V, typename I, typename P>
__device__ __forceinline__ void entry_xgw_fetch(const SpmvParams<VS, V, I, P>& p, int64_t t,
                                                XgwFetch& g) {
  constexpr int T = E * kBlock;
  const int tid = threadIdx.x;
  g.e0 = p.tile_entry[t];
  int64_t e1 = p.tile_entry[t + 1];
  e1 = e1 > p.nnz ? p.nnz : e1;
  const int64_t span = e1 - g.e0;
  g.len = span <= 0 ? 0 : (span < T ? (int)span : T);  // as entry_fetch clamps it
  if (g.len == 0) return;  // uniform: an empty tile reads nothing
  const int64_t rmax = p.xg_n_runs - 1;
  int64_t q0 = p.xg_tile_run[t], q1 = p.xg_tile_run[t + 1];
  q0 = q0 < 0 ? 0 : (q0 > rmax ? rmax : q0);
  q1 = q1 <= q0 ? q0 + 1 : (q1 > rmax + 1 ? rmax + 1 : q1);
  g.q0 = q0;
  g.nr = (int)(q1 - q0 < T ? q1 - q0 : T);
  g.b0 = g.e0 >> 6;
  g.nb = (int)(((g.e0 + g.len - 1) >> 6) - g.b0 + 1);  // <= T / 64 + 1
  const int bi = tid < g.nb ? tid : g.nb - 1;
  g.br = p.xg_batch_run[g.b0 + bi];
  g.bm = p.xg_batch_mask[g.b0 + bi];
  g.ro = p.xg_run_off[q0 + (tid < g.nr ? tid : g.nr - 1)];
}

// Stores the run table, then issues every position's x load (v); called
// before the tile's entry loads, so those and these are in flight together.
template <int E, typename VS, typename V, typename I, typename P>
__device__ __forceinline__ void entry_xgw_issue(const SpmvParams<VS, V, I, P>& p,
                                                const XgwFetch& g, const XgwLds<E * kBlock>& L,
                                                V (&v)[E]) {
  const int tid = threadIdx.x;
  if (g.len == 0) return;  // uniform: an empty tile reads nothing
  if (tid < g.nb) {
    L.run[tid] = g.br;
    L.mask[tid] = g.bm;
  }
  if (tid < g.nr) L.off[tid] = g.ro;
  // a tile with more than kBlock runs (windows of few columns): the rest
  for (int i = kBlock + tid; i < g.nr; i += kBlock) L.off[i] = p.xg_run_off[g.q0 + i];
  lds_barrier();
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int o = u * kBlock + tid;
    const int64_t d = g.e0 + (o < g.len ? o : 0);
    const int b = (int)((d >> 6) - g.b0);
    const uint64_t upto = (2ull << (d & 63)) - 2;  // bits 1 .. d % 64
    int64_t r = (int64_t)L.run[b] + __builtin_popcountll(L.mask[b] & upto) - g.q0;
    r = r < 0 ? 0 : (r >= g.nr ? g.nr - 1 : r);
    int64_t src = d + L.off[r];
    src = src < 0 ? 0 : (src >= p.nnz ? p.nnz - 1 : src);
    v[u] = p.x[src];
  }
}

// After every read of the run table, which xt may overlay: the x values to LDS.
template <int E, typename V>
__device__ __forceinline__ void entry_xgw_store(const XgwFetch& g, const V (&v)[E], V* xt) {
  const int tid = threadIdx.x;
  lds_barrier();
  if (g.len == 0) return;
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int o = u * kBlock + tid;
    if (o < g.len) xt[o] = v[u];
  }
}

template <int E, typename I, typename C, typename V>
__device__ __forceinline__ void entry_gather_xg(const EntryTile<I, C, E>& f, const V* xt,
                                                C (&xv)[E]) {
  constexpr int T = E * kBlock;
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int q = (int)f.col[u];
    xv[u] = load_as<C>(xt, q < T ? q : T - 1);
  }
}

// Products to LDS in entry order, and the tile's row boundaries (entry
// offsets clamped to [0, len]) of its first batch of owned rows.
template <int E, typename I, typename C>
__device__ __forceinline__ void entry_stage(const EntryTile<I, C, E>& f, const C (&xv)[E], C* prod,
                                            int32_t* bnd) {
  const int tid = threadIdx.x;
  if (f.vec) {
#pragma unroll
    for (int g = 0; g < E / 4; ++g)
#pragma unroll
      for (int k = 0; k < 4; ++k) prod[(g * kBlock + tid) * 4 + k] = mulc(f.a[4 * g + k], xv[4 * g + k]);
  } else {
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int o = u * kBlock + tid;
      if (o < f.len) prod[o] = mulc(f.a[u], xv[u]);
    }
  }
#pragma unroll
  for (int k = 0; k < EntryTile<I, C, E>::kPv; ++k) {
    const int i = k * kBlock + tid;
    if (i <= f.nb0) {
      const int64_t v = f.pv[k] - f.e0;
      bnd[i] = (int)(v < 0 ? 0 : (v > f.len ? f.len : v));
      // whether the last owned row runs past the tile, from the pointer
      // already loaded (-1: the owned rows exceed one batch, decided later)
      if (i == f.nb0)
        bnd[kStraddleSlot] = f.r1 - f.r0 > kSegRows ? -1 : (f.r1 > f.r0 && v > f.len ? 1 : 0);
    }
  }
  if (tid == 0) bnd[kLongCountSlot] = 0;
}

// After the LDS stage and a barrier: the head carry (entries of a row owned
// by an earlier tile), then every owned row summed in fp64 in entry order,
// a row running past the tile going to the tail carry.
template <typename A, int E, typename VS, typename V, typename I, typename P, typename C, typename S>
__device__ __forceinline__ void entry_rows(const SpmvParams<VS, V, I, P>& p,
                                           const EntryTile<I, C, E>& f, const C* prod,
                                           int32_t* bnd, S* red, S* carries) {
  const int tid = threadIdx.x, w = tid / kWave, lane = tid % kWave;
  const int64_t t = f.t, e0 = f.e0, r0 = f.r0, r1 = f.r1;
  const int len = f.len;
  const int sflag = bnd[kStraddleSlot];
  const bool straddle = sflag >= 0 ? sflag == 1 : (r1 > r0 && (int64_t)p.ptr[r1] > e0 + len);
  // head: entries [0, bnd[0]) belong to a row owned by an earlier tile
  const int head = bnd[0];
  if (head > 0 && head <= 4 * kWave) {
    if (w == 0) {
      Plain<S> acc;
      acc.init();
      for (int j = lane; j < head; j += kWave) acc.add_value(to_sum<S>(prod[j]));
      group_reduce<kWave>(acc);
      if (lane == 0) carries[2 * t] = acc.value();
    }
  } else if (head > 0) {
    Plain<S> acc;
    acc.init();
    for (int j = tid; j < head; j += kBlock) acc.add_value(to_sum<S>(prod[j]));
    group_reduce<kWave>(acc);
    if (lane == 0) red[w] = acc.value();
    __syncthreads();
    if (tid == 0) {
      S tot = red[0];
      for (int k = 1; k < kWavesPerBlock; ++k) tot = add(tot, red[k]);
      carries[2 * t] = tot;
    }
  } else if (tid == 0) {
    carries[2 * t] = zero<S>();
  }
  for (int64_t rb = r0; rb < r1; rb += kSegRows) {
    const int nb = (int)(r1 - rb < kSegRows ? r1 - rb : kSegRows);
    if (rb != r0) {  // later batches (tiles owning more than kSegRows rows)
      __syncthreads();
      for (int i = tid; i <= nb; i += kBlock) {
        const int64_t v = (int64_t)p.ptr[rb + i] - e0;
        bnd[i] = (int)(v < 0 ? 0 : (v > len ? len : v));
      }
      if (tid == 0) bnd[kLongCountSlot] = 0;
      __syncthreads();
    }
    // short rows: a thread each; longer rows are listed for the wavefronts
    // (at most len / (kThreadRowMax + 1) < kLongRowsMax of them)
    for (int i = tid; i < nb; i += kBlock) {
      const int s = bnd[i], e = bnd[i + 1];
      if (e - s > kThreadRowMax) {
        const int slot = atomicAdd(&bnd[kLongCountSlot], 1);
        if (slot < kLongRowsMax) bnd[kLongList + slot] = i;
        continue;
      }
      S acc = zero<S>();
      for (int j = s; j < e; ++j) acc = add(acc, to_sum<S>(prod[j]));
      const int64_t row = rb + i;
      entry_finish<A>(p, row, acc, straddle && row == r1 - 1, carries, t);
    }
    __syncthreads();
    int n_long = bnd[kLongCountSlot];
    n_long = n_long < kLongRowsMax ? n_long : kLongRowsMax;
    for (int q = w; q < n_long; q += kWavesPerBlock) {
      const int i = bnd[kLongList + q];
      const int s = bnd[i], e = bnd[i + 1];
      Plain<S> acc;
      acc.init();
      for (int j = s + lane; j < e; j += kWave) acc.add_value(to_sum<S>(prod[j]));
      group_reduce<kWave>(acc);
      const int64_t row = rb + i;
      if (lane == 0) entry_finish<A>(p, row, acc.value(), straddle && row == r1 - 1, carries, t);
    }
  }
}

// LDS carve-up shared by the entry kernels: products, row boundaries, the
// per-wavefront reduction slots.
template <typename C, typename S, int T> struct EntryLds {
  C* prod;
  int32_t* bnd;
  S* red;
  __device__ __forceinline__ explicit EntryLds(char* smem)
      : prod(reinterpret_cast<C*>(smem)),
        bnd(reinterpret_cast<int32_t*>(smem + sizeof(C) * T)),
        red(reinterpret_cast<S*>(smem + sizeof(C) * T + sizeof(int32_t) * kBndSlots)) {}
};

// One tile per block (block b takes tile b, or with block groups the b / G-th
// tile of group b % G).
template <typename A, int E, bool VEC, int XG, typename VS, typename V, typename I, typename P>
__global__ __launch_bounds__(kBlock) void spmv_entry_kernel(SpmvParams<VS, V, I, P> p) {
  using C = typename A::compute_t;
  using S = typename TileSumOf<A>::T;
  extern __shared__ __align__(16) char smem[];
  const EntryLds<C, S, E * kBlock> lds(smem);
  S* carries = reinterpret_cast<S*>(p.partials);
  int64_t t = blockIdx.x;
  if (p.n_groups > 0) {  // block b serves group b % G, in order within the group
    const int g = (int)(t % p.n_groups);
    t = p.group_tiles[g] + t / p.n_groups;
    if (t >= p.group_tiles[g + 1] || t >= p.n_tiles) return;
  }
  EntryTile<I, C, E> f;
  V* xt = reinterpret_cast<V*>(smem + entry_xg_offset<C, V>(E * kBlock));
  XgwFetch g;
  V xw[XG == 2 ? E : 1];
  if constexpr (XG == 2) {  // run table, then the tile's x loads, then its entry loads
    const XgwLds<E * kBlock> L(reinterpret_cast<char*>(lds.bnd), smem);
    entry_xgw_fetch<E>(p, t, g);
    entry_xgw_issue<E>(p, g, L, xw);
  }
  entry_fetch<E, VEC>(p, t, f);
  C xv[E];
  if constexpr (XG != 0) {
    if constexpr (XG == 2) {
      entry_xgw_store<E>(g, xw, xt);
    } else
      entry_stage_xg(p, f, xt);
    __syncthreads();
    entry_gather_xg(f, xt, xv);
    if constexpr (kXgAliased<C, V>) __syncthreads();  // every read of xt before the products
  } else {
    entry_gather(p, f, xv);
  }
  entry_stage(f, xv, lds.prod, lds.bnd);
  __syncthreads();
  entry_rows<A>(p, f, lds.prod, lds.bnd, lds.red, carries);
}

// Rows cut by a tile boundary: thread t takes tile t's last row when it runs
// past the tile and adds the heads of the following tiles in order until the
// tile that owns the next row (or the last tile).
template <typename A, typename VS, typename V, typename I, typename P>
__global__ __launch_bounds__(kBlock) void spmv_entry_fixup_kernel(SpmvParams<VS, V, I, P> p) {
  using S = typename TileSumOf<A>::T;
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t + 1 >= p.n_tiles) return;  // the last tile ends at nnz: no row runs past it
  const S* carries = reinterpret_cast<const S*>(p.partials);
  // the following tiles kWalk at a time: their heads and row starts are
  // loaded together (indices clamped to the last tile, so every load is in
  // bounds), then added in tile order up to the tile that owns the next row
  // -- the same sum as one tile at a time, without a dependent load per tile
  constexpr int kWalk = 8;
  const int64_t last = p.n_tiles - 1;
  S head[kWalk];
  int64_t ts[kWalk + 1];
  auto load_walk = [&](int64_t u0) {
#pragma unroll
    for (int k = 0; k < kWalk; ++k) {
      const int64_t u = u0 + k < last ? u0 + k : last;
      head[k] = carries[2 * u];
      ts[k] = p.tile_start[u];
    }
    ts[kWalk] = p.tile_start[u0 + kWalk < p.n_tiles ? u0 + kWalk : p.n_tiles];
  };
  // The first walk depends only on t, so it is loaded with the straddle
  // test's own operands rather than after them: two dependent round trips
  // (tile bounds, then the row pointer) instead of three for the common
  // one-tile case.
  load_walk(t + 1);
  S acc = carries[2 * t + 1];
  const int64_t r0 = p.tile_start[t], r1 = ts[0];
  const int64_t e1 = p.tile_entry[t + 1];
  if (r1 <= r0 || r1 > p.n_rows) return;
  if ((int64_t)p.ptr[r1] <= e1) return;
  bool done = false;
  for (int64_t u0 = t + 1;;) {
#pragma unroll
    for (int k = 0; k < kWalk; ++k) {
      if (done || u0 + k >= p.n_tiles) break;
      acc = add(acc, head[k]);
      const int64_t next = k + 1 < kWalk ? ts[k + 1] : ts[kWalk];
      if (u0 + k == last || next > ts[k]) done = true;
    }
    u0 += kWalk;
    if (done || u0 >= p.n_tiles) break;
    load_walk(u0);
  }
  entry_write(p, r1 - 1, acc);
}

template <typename A, int E, bool VEC, int XG, typename VS, typename V, typename I, typename P>
void spmv_launch_entry_e(const SpmvParams<VS, V, I, P>& p, int64_t grid, hipStream_t st) {
  using C = typename A::compute_t;
  const size_t lds = entry_kernel_lds<C, V>(XG, E * kBlock);
  spmv_entry_kernel<A, E, VEC, XG><<<dim3((unsigned)grid), dim3(kBlock), lds, st>>>(p);
}

// Tile-major (XG 1) or window-major (XG 2, p.xg_run_off set) gathered x.
template <typename A, int E, bool VEC, typename VS, typename V, typename I, typename P>
void spmv_launch_entry_g(const SpmvParams<VS, V, I, P>& p, int64_t grid, hipStream_t st) {
  if (p.xg_run_off) spmv_launch_entry_e<A, E, VEC, 2>(p, grid, st);
  else spmv_launch_entry_e<A, E, VEC, 1>(p, grid, st);
}

template <typename A, int E, bool VEC, bool XG, typename VS, typename V, typename I, typename P>
void spmv_launch_entry_x(const SpmvParams<VS, V, I, P>& p, int64_t grid, hipStream_t st) {
  if constexpr (XG) spmv_launch_entry_g<A, E, VEC>(p, grid, st);
  else spmv_launch_entry_e<A, E, VEC, 0>(p, grid, st);
}

// Whether the col / value arrays allow the 16-byte loads of the entry role.
template <typename VS, typename V, typename I, typename P>
bool entry_vec_ok(const SpmvParams<VS, V, I, P>& p) {
  if constexpr (!VS::kVec) {
    return false;
  } else {
    return ((uintptr_t)p.idx % 16) == 0 && p.vals.vec_aligned();
  }
}

// Launches the entry-tile kernel and its fix-up.  HOT builds tiles of 1024,
// 2048 and 4096 entries with and without vector loads; others 2048, scalar.
// XG: the gathered-x form (x holds per-entry x values, p.idx is x_perm).
template <typename A, bool HOT, bool XG = false, typename VS, typename V, typename I, typename P>
int spmv_launch_entry(const SpmvParams<VS, V, I, P>& p, int64_t grid, hipStream_t st,
                      const char* who) {
  using C = typename A::compute_t;
  const size_t lds = entry_kernel_lds<C, V>(XG ? (p.xg_run_off ? 2 : 1) : 0, p.tile_nnz);
  if (lds > 65536)
    return fail(FS_ERR_VALUE, "%s: entry tiles of %lld entries need more than 64 KB of LDS", who,
                (long long)p.tile_nnz);
  bool vec = entry_vec_ok(p);
  if constexpr (XG) vec = vec && ((uintptr_t)p.x % 16) == 0;
  bool done = false;
  if constexpr (HOT) {
    switch (p.tile_nnz) {
      case 1024: vec ? spmv_launch_entry_x<A, 4, true, XG>(p, grid, st) : spmv_launch_entry_x<A, 4, false, XG>(p, grid, st); done = true; break;
      case 2048: vec ? spmv_launch_entry_x<A, 8, true, XG>(p, grid, st) : spmv_launch_entry_x<A, 8, false, XG>(p, grid, st); done = true; break;
      case 4096:
        vec ? spmv_launch_entry_x<A, 16, true, XG>(p, grid, st) : spmv_launch_entry_x<A, 16, false, XG>(p, grid, st);
        done = true;
        break;
      default: break;
    }
  } else if (p.tile_nnz == 2048) {
    spmv_launch_entry_x<A, 8, false, XG>(p, grid, st);
    done = true;
  }
  if (!done)
    return fail(FS_ERR_VALUE, "%s: entry tiles of %lld entries are not built for this type", who,
                (long long)p.tile_nnz);
  int rc = check_launch(who);
  if (rc || p.n_tiles < 2) return rc;
  const int64_t blocks = (p.n_tiles + kBlock - 1) / kBlock;
  spmv_entry_fixup_kernel<A><<<dim3((unsigned)blocks), dim3(kBlock), 0, st>>>(p);
  return check_launch(who);
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
