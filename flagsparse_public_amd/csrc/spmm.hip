// spmm.hip -- CSR SpMM for gfx950: C = alpha * A B + beta * C, row-major B and C.
//
// A group of G lanes owns one row and a tile of G*VW dense columns; each lane
// owns VW consecutive columns and gathers them from the B row of every entry
// with one vector load (VW * element size bytes, e.g. 16 B for fp32 x4), so a
// group reads one contiguous B-row segment per entry.  U entries are kept in
// flight per lane.  Rows longer than short_max are cut into chunks reduced by
// whole wavefronts (64/G groups interleave over the chunk's entries and are
// combined with xor exchanges); a row spanning several chunks stores one
// partial row per chunk, and spmm_fixup_kernel (a second, small launch) sums
// them in chunk order.
//
// Replaces the reference's ALG1 row x column-tile kernel (spmm_csr.py:92-134,
// 525-599) and the bucketed opt / alg2 kernels (spmm_csr.py:602-919,
// spmm_csr_opt_alg2.py:338-629).
#include <type_traits>

#include "fs_host.hpp"

namespace fs {

constexpr int kMmBlock = 256;
constexpr int kMmWaves = kMmBlock / kWave;

template <typename V, typename I, typename P>
struct SpmmParams {
  const P* ptr;
  const I* idx;
  const V* val;
  const V* B;
  V* C;
  int64_t ldb, ldc, n_dense;
  int64_t n_rows;
  int64_t n_tiles;  // column tiles of G*VW columns
  double alpha, beta;
  int64_t short_max;
  int64_t n_chunk_blocks;
  int64_t n_chunks;
  int64_t chunk_nnz;
  const int64_t* long_rows;
  const int64_t* long_chunk_ptr;
  const int32_t* chunk_long;
  void* partials;     // [n_chunks * n_dense] accumulators
  int xcd_remap;
};

template <typename A> struct MmChunkAcc { using T = A; };
template <> struct MmChunkAcc<Plain<float>> { using T = Comp; };

// True xor exchange (lane ^ K): quad_perm for 1 and 2, ds_swizzle for 4..16,
// ds_bpermute for 32.  Used to combine lane groups at equal positions.
template <int K> __device__ __forceinline__ int xor_i32(int v) {
  if constexpr (K == 1) return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
  else if constexpr (K == 2) return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);
  else if constexpr (K == 4 || K == 8 || K == 16) return __builtin_amdgcn_ds_swizzle(v, 0x1F | (K << 10));
  else {
    static_assert(K == 32, "bad xor distance");
    return __builtin_amdgcn_ds_bpermute((__lane_id() ^ 32) << 2, v);
  }
}
template <int K> __device__ __forceinline__ float xor_x(float v) { return __int_as_float(xor_i32<K>(__float_as_int(v))); }
template <int K> __device__ __forceinline__ double xor_x(double v) {
  int64_t b = __double_as_longlong(v);
  int lo = xor_i32<K>((int)(b & 0xffffffff));
  int hi = xor_i32<K>((int)(b >> 32));
  return __longlong_as_double(((int64_t)(uint32_t)lo) | ((int64_t)hi << 32));
}
template <int K, typename T> __device__ __forceinline__ cplx<T> xor_x(cplx<T> v) { return {xor_x<K>(v.re), xor_x<K>(v.im)}; }
template <int K, typename T> __device__ __forceinline__ Plain<T> xor_x(const Plain<T>& a) { Plain<T> o; o.s = xor_x<K>(a.s); return o; }
template <int K> __device__ __forceinline__ Comp xor_x(const Comp& a) { Comp o; o.s = xor_x<K>(a.s); o.c = xor_x<K>(a.c); return o; }

// Sum accumulators over lane groups of width G inside a wavefront (positions
// kept): after this every lane holds the sum over the 64/G groups.
template <int G, typename A> __device__ __forceinline__ void cross_group_reduce(A& a) {
  if constexpr (G <= 1) { A o = xor_x<1>(a); a.merge(o); }
  if constexpr (G <= 2) { A o = xor_x<2>(a); a.merge(o); }
  if constexpr (G <= 4) { A o = xor_x<4>(a); a.merge(o); }
  if constexpr (G <= 8) { A o = xor_x<8>(a); a.merge(o); }
  if constexpr (G <= 16) { A o = xor_x<16>(a); a.merge(o); }
  if constexpr (G <= 32) { A o = xor_x<32>(a); a.merge(o); }
}

// Load VW consecutive stored values starting at element `off` as compute type.
template <typename C, int VW, typename S>
__device__ __forceinline__ void load_vec(C (&out)[VW], const S* __restrict__ p, int64_t off) {
  if constexpr (std::is_same<S, float>::value && VW == 4) {
    const float4 v = *reinterpret_cast<const float4*>(p + off);
    out[0] = (C)v.x; out[1] = (C)v.y; out[2] = (C)v.z; out[3] = (C)v.w;
  } else if constexpr (std::is_same<S, float>::value && VW == 2) {
    const float2 v = *reinterpret_cast<const float2*>(p + off);
    out[0] = (C)v.x; out[1] = (C)v.y;
  } else if constexpr (std::is_same<S, double>::value && VW == 2) {
    const double2 v = *reinterpret_cast<const double2*>(p + off);
    out[0] = (C)v.x; out[1] = (C)v.y;
  } else {
#pragma unroll
    for (int k = 0; k < VW; ++k) out[k] = load_as<C>(p, off + k);
  }
}

// B rows of a masked batch: b[u] = B[col[u] * ldb + n0 .. + VW) for the u
// whose bit is set in `ok` (others zero).  When the compute type differs from
// the stored one (fp32 B accumulated in fp64), the rows are loaded raw and kept
// opaque (empty asm) until every load of the batch is issued, then widened: a
// widening placed inside a masked load's branch brings a wait for that load
// with it and serialises the batch.
template <int VW, int U, typename C, typename V>
__device__ __forceinline__ void gather_masked(C (&b)[U][VW], const V* __restrict__ B, int64_t ldb,
                                              const int64_t (&col)[U], int64_t n0, unsigned ok) {
  if constexpr (std::is_same<C, V>::value || !std::is_arithmetic<V>::value) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if ((ok >> u) & 1u) {
        load_vec<C, VW>(b[u], B, col[u] * ldb + n0);
      } else {
#pragma unroll
        for (int k = 0; k < VW; ++k) b[u][k] = zero<C>();
      }
    }
  } else {
    V r[U][VW];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if ((ok >> u) & 1u) {
        load_vec<V, VW>(r[u], B, col[u] * ldb + n0);
      } else {
#pragma unroll
        for (int k = 0; k < VW; ++k) r[u][k] = V(0);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int k = 0; k < VW; ++k) {
        asm volatile("" : "+v"(r[u][k]));
        b[u][k] = (C)r[u][k];
      }
  }
}

// C[off .. off + VW) = alpha * acc + beta * C (C read only when beta is
// nonzero): write_out's arithmetic, with one vector load / store where
// load_vec has a vector form (the launch keeps C aligned whenever VW > 1) and
// the row's lane group is at least 32 lanes wide.  Measured on the 1M-row
// banded matrix: vector stores took N = 128 (32 lanes) from 798 to 776 us,
// but were ~2 us slower than VW scalar stores at N = 32 / 64 (8 / 16 lanes).
template <int G, int VW, typename V, typename A>
__device__ __forceinline__ void write_vec(V* __restrict__ C, int64_t off, const A (&acc)[VW],
                                          double alpha, double beta) {
  constexpr bool kVec = G >= 32 && ((std::is_same<V, float>::value && (VW == 4 || VW == 2)) ||
                                    (std::is_same<V, double>::value && VW == 2));
  if constexpr (kVec) {
    using T = std::decay_t<decltype(acc[0].value())>;
    T v[VW];
#pragma unroll
    for (int k = 0; k < VW; ++k) v[k] = alpha == 1.0 ? acc[k].value() : scale(acc[k].value(), alpha);
    if (beta != 0.0) {
      T old[VW];
      load_vec<T, VW>(old, C, off);
#pragma unroll
      for (int k = 0; k < VW; ++k) v[k] = add(v[k], scale(old[k], beta));
    }
    if constexpr (std::is_same<V, float>::value && VW == 4)
      *reinterpret_cast<float4*>(C + off) = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
    else if constexpr (std::is_same<V, float>::value)
      *reinterpret_cast<float2*>(C + off) = make_float2((float)v[0], (float)v[1]);
    else
      *reinterpret_cast<double2*>(C + off) = make_double2((double)v[0], (double)v[1]);
  } else {
#pragma unroll
    for (int k = 0; k < VW; ++k) write_out(C, off + k, acc[k].value(), alpha, beta);
  }
}

// Accumulate entries j = s + first, s + first + step, ... < e into acc[VW] for
// the lane's columns n0 .. n0+VW-1.
template <int VW, int U, typename A, typename V, typename I>
__device__ __forceinline__ void mm_accumulate(A (&acc)[VW], const I* __restrict__ idx,
                                              const V* __restrict__ val,
                                              const V* __restrict__ B, int64_t ldb,
                                              int64_t n0, int64_t s, int64_t e, int64_t step) {
  using C = typename A::compute_t;
  int64_t j = s;
  for (; j + (int64_t)(U - 1) * step < e; j += (int64_t)U * step) {
    int64_t col[U];
    C a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      col[u] = ldg_index(idx, j + u * step);
      a[u] = load_as<C>(val, j + u * step);
    }
    C b[U][VW];
#pragma unroll
    for (int u = 0; u < U; ++u) load_vec<C, VW>(b[u], B, col[u] * ldb + n0);
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int k = 0; k < VW; ++k) acc[k].add_prod(a[u], b[u][k]);
  }
  if (j < e) {
    // the last, partial batch: all its loads in flight at once, entries past
    // e masked (their products are not added)
    I ci[U];
    C a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool ok = j + u * step < e;
      ci[u] = ok ? idx[j + u * step] : I(0);
      a[u] = ok ? load_as<C>(val, j + u * step) : zero<C>();
    }
    // the empty asm keeps each index's widening (and the wait for its load)
    // out of the masked load's branch, so the batch's loads are all issued
    // before the first wait
    int64_t col[U];
    unsigned ok = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      asm volatile("" : "+v"(ci[u]));
      col[u] = (int64_t)ci[u];
      ok |= (unsigned)(j + u * step < e) << u;
    }
    C b[U][VW];
    gather_masked<VW, U>(b, B, ldb, col, n0, ok);
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (j + u * step < e)
#pragma unroll
        for (int k = 0; k < VW; ++k) acc[k].add_prod(a[u], b[u][k]);
  }
}

// A short row's entries, loaded G at a time by the G lanes of its group (lane
// g loads entry j0 + g: one coalesced col / value load per G entries instead
// of one broadcast load per entry, which cuts the L2 requests of the entry
// stream), parked in the group's LDS slots and read back U at a time.  Lanes
// past the last dense column (colok false) still load their share.  The
// products are added in entry order, as in mm_accumulate.
// F fills per round: each lane loads F entries (F * G slots per group), so a
// row of up to F * G entries needs one fill round trip instead of F.
template <int G> constexpr int kMmFill = G <= 8 ? 2 : 1;

template <int VW, int U, int G, typename A, typename V>
__device__ __forceinline__ void mm_accumulate_group(A (&acc)[VW], int32_t* slot_col,
                                                    typename A::compute_t* slot_val,
                                                    const int32_t* __restrict__ idx,
                                                    const V* __restrict__ val,
                                                    const V* __restrict__ B, int64_t ldb,
                                                    int64_t n0, bool colok, int64_t s,
                                                    int64_t e, int g) {
  using C = typename A::compute_t;
  constexpr int F = kMmFill<G>;
  static_assert(G % U == 0, "batches of U entries tile the group's G slots");
  for (int64_t j0 = s; j0 < e; j0 += G * F) {
    if constexpr (F == 1) {
      const int64_t jj = j0 + g;
      slot_col[g] = jj < e ? idx[jj] : 0;
      slot_val[g] = jj < e ? load_as<C>(val, jj) : zero<C>();
    } else {
      // unconditional loads (past e: entry j0 again), then the LDS stores,
      // so the F loads of a lane are in flight together
      int32_t fc[F];
      V fv[F];
#pragma unroll
      for (int f = 0; f < F; ++f) {
        const int64_t jj = j0 + f * G + g;
        const int64_t q = jj < e ? jj : j0;
        fc[f] = idx[q];
        fv[f] = val[q];
      }
#pragma unroll
      for (int f = 0; f < F; ++f) {
        slot_col[f * G + g] = fc[f];
        slot_val[f * G + g] = (C)fv[f];
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    const int cnt = e - j0 < G * F ? (int)(e - j0) : G * F;
    if (colok) {
      for (int u0 = 0; u0 < cnt; u0 += U) {
        int64_t col[U];
        C a[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          // LDS reads are unconditional (a slot past cnt re-reads slot u0):
          // under a per-lane branch each would be waited for in turn
          const int uu = u0 + u < cnt ? u0 + u : u0;
          col[u] = (int64_t)slot_col[uu];
          a[u] = slot_val[uu];
        }
        unsigned ok = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) ok |= (unsigned)(u0 + u < cnt) << u;
        C b[U][VW];
        gather_masked<VW, U>(b, B, ldb, col, n0, ok);
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (u0 + u < cnt)
#pragma unroll
            for (int k = 0; k < VW; ++k) acc[k].add_prod(a[u], b[u][k]);
      }
    }
    // the slots are rewritten by the next batch
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
}

template <typename V, typename A, typename I, typename P, int G, int VW, int U>
__device__ __forceinline__ void spmm_chunk_role(const SpmmParams<V, I, P>& p, int64_t bid) {
  using CA = typename MmChunkAcc<A>::T;
  constexpr int kGroups = kWave / G;
  const int lane = threadIdx.x % kWave;
  const int64_t item = bid * kMmWaves + threadIdx.x / kWave;
  const int64_t c = item / p.n_tiles, tile = item % p.n_tiles;
  if (c >= p.n_chunks) return;
  const int grp = lane / G, g = lane % G;
  const int64_t n0 = tile * (G * VW) + (int64_t)g * VW;
  const bool colok = n0 < p.n_dense;
  const int64_t li = p.chunk_long[c];
  const int64_t row = p.long_rows[li];
  const int64_t c0 = p.long_chunk_ptr[li];
  const int64_t nc = p.long_chunk_ptr[li + 1] - c0;
  const int64_t rs = (int64_t)p.ptr[row], re = (int64_t)p.ptr[row + 1];
  const int64_t s = rs + (c - c0) * p.chunk_nnz;
  const int64_t e = s + p.chunk_nnz < re ? s + p.chunk_nnz : re;
  CA acc[VW];
#pragma unroll
  for (int k = 0; k < VW; ++k) acc[k].init();
  if (colok) mm_accumulate<VW, U>(acc, p.idx, p.val, p.B, p.ldb, n0, s + grp, e, kGroups);
#pragma unroll
  for (int k = 0; k < VW; ++k) cross_group_reduce<G>(acc[k]);
  if (nc == 1) {
    if (grp == 0 && colok) write_vec<G, VW>(p.C, row * p.ldc + n0, acc, p.alpha, p.beta);
    return;
  }
  CA* parts = reinterpret_cast<CA*>(p.partials);  // summed by spmm_fixup_kernel
  if (grp == 0 && colok)
#pragma unroll
    for (int k = 0; k < VW; ++k) parts[c * p.n_dense + n0 + k] = acc[k];
}

// Second pass for rows cut into several chunks: one wavefront per long row;
// lane n sums column n's chunk partials in chunk order (then n + 64, ...).
template <typename A, typename V, typename I, typename P>
__global__ __launch_bounds__(kMmBlock) void spmm_fixup_kernel(SpmmParams<V, I, P> p, int64_t n_long) {
  using CA = typename MmChunkAcc<A>::T;
  const int lane = threadIdx.x % kWave;
  const int64_t li = (int64_t)blockIdx.x * kMmWaves + threadIdx.x / kWave;
  if (li >= n_long) return;
  const int64_t c0 = p.long_chunk_ptr[li];
  const int64_t nc = p.long_chunk_ptr[li + 1] - c0;
  if (nc <= 1) return;
  const int64_t row = p.long_rows[li];
  const CA* parts = reinterpret_cast<const CA*>(p.partials);
  for (int64_t n = lane; n < p.n_dense; n += kWave) {
    CA tot;
    tot.init();
    for (int64_t k = 0; k < nc; ++k) tot.merge(parts[(c0 + k) * p.n_dense + n]);
    write_out(p.C, row * p.ldc + n, tot.value(), p.alpha, p.beta);
  }
}

template <typename V, typename I>
constexpr bool kHotSpmm = (std::is_same<V, float>::value || std::is_same<V, double>::value) &&
                          std::is_same<I, int32_t>::value;

// Short rows load their entries group-cooperatively for the fp32 / fp64,
// int32-column kernels whose lane groups hold whole batches.
template <typename V, typename I, int G, int U>
constexpr bool kGroupLoads = kHotSpmm<V, I> && G >= U && G % U == 0;

// Lane groups narrower than a batch (G < U: few dense columns) take the
// wavefront's 64 / G consecutive rows together: their entries are
// contiguous, so the wavefront copies them into its LDS window with coalesced
// loads (64 consecutive entries per instruction) and each group reads its
// row's columns and values from there.  A wavefront whose rows hold more than
// kMmStageMax entries (a long row among them) loads per group instead.
template <typename V, typename I, int G, int U>
constexpr bool kWaveStaged = kHotSpmm<V, I> && G < U;

constexpr int kMmStage = 512;          // entries per wavefront window
constexpr int kMmStageMax = 8 * kMmStage;

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

template <typename A, int G, int VW, int U, typename V, typename I, typename P>
__device__ __forceinline__ void spmm_staged_rows(const SpmmParams<V, I, P>& p, int64_t rb,
                                                 int64_t row, int64_t n0) {
  using C = typename A::compute_t;
  __shared__ int32_t st_col[kMmWaves][kMmStage];
  __shared__ C st_val[kMmWaves][kMmStage];
  const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
  const int64_t wrow0 = rb * (kMmBlock / G) + (int64_t)wv * (kWave / G);
  if (wrow0 >= p.n_rows) return;  // the whole wavefront
  const int64_t wrow1 = wrow0 + kWave / G < p.n_rows ? wrow0 + kWave / G : p.n_rows;
  const int64_t wb = (int64_t)p.ptr[wrow0], we = (int64_t)p.ptr[wrow1];
  const bool live = row < wrow1;
  const int64_t rs = live ? (int64_t)p.ptr[row] : we, re = live ? (int64_t)p.ptr[row + 1] : we;
  const bool mine = live && re - rs <= p.short_max && n0 < p.n_dense;  // longer: chunked
  A acc[VW];
#pragma unroll
  for (int k = 0; k < VW; ++k) acc[k].init();
  if (we - wb > kMmStageMax) {
    if (!mine) return;
    mm_accumulate<VW, U>(acc, p.idx, p.val, p.B, p.ldb, n0, rs, re, 1);
  } else {
    int32_t* sc = st_col[wv];
    C* sv = st_val[wv];
    for (int64_t w0 = wb; w0 < we; w0 += kMmStage) {
      const int64_t cnt = we - w0 < kMmStage ? we - w0 : kMmStage;
      // all of the window's loads first, then its LDS stores: a load whose
      // store sits in the same per-lane branch is waited for there, one slice
      // of the window at a time.  Slices wholly past cnt are skipped by the
      // wavefront (cnt is uniform); lanes past cnt re-read entry w0.
      constexpr int kPer = kMmStage / kWave;
      const int cntu = __builtin_amdgcn_readfirstlane((int)cnt);
      int32_t lc[kPer];
      V lv[kPer];
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        if (k * kWave < cntu) {
          const int i = k * kWave + lane;
          const int64_t jj = w0 + (i < cntu ? i : 0);
          lc[k] = (int32_t)p.idx[jj];
          lv[k] = p.val[jj];
        } else {
          lc[k] = 0;
          lv[k] = V(0);
        }
      }
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        asm volatile("" : "+v"(lc[k]));
        asm volatile("" : "+v"(lv[k]));
      }
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const int i = k * kWave + lane;
        if (i < cntu) {
          sc[i] = lc[k];
          sv[i] = (C)lv[k];
        }
      }
      wave_lds_sync();
      const int64_t lo = rs > w0 ? rs : w0;
      const int64_t hi = re < w0 + cnt ? re : w0 + cnt;
      if (mine) {
        for (int64_t j = lo; j < hi; j += U) {
          int64_t col[U];
          C a[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            // unconditional LDS reads (past hi: entry j again), as above
            const int64_t jj = j + u < hi ? j + u : j;
            col[u] = (int64_t)sc[jj - w0];
            a[u] = sv[jj - w0];
          }
          unsigned ok = 0;
#pragma unroll
          for (int u = 0; u < U; ++u) ok |= (unsigned)(j + u < hi) << u;
          C b[U][VW];
          gather_masked<VW, U>(b, p.B, p.ldb, col, n0, ok);
#pragma unroll
          for (int u = 0; u < U; ++u)
            if (j + u < hi)
#pragma unroll
              for (int k = 0; k < VW; ++k) acc[k].add_prod(a[u], b[u][k]);
        }
      }
      wave_lds_sync();  // the window is rewritten next
    }
    if (!mine) return;
  }
  write_vec<G, VW>(p.C, row * p.ldc + n0, acc, p.alpha, p.beta);
}

template <typename V, typename A, typename I, typename P, int G, int VW, int U>
__global__ __launch_bounds__(kMmBlock) void spmm_csr_kernel(SpmmParams<V, I, P> p) {
  const int64_t bid = blockIdx.x;
  if (bid < p.n_chunk_blocks) {
    spmm_chunk_role<V, A, I, P, G, VW, U>(p, bid);
    return;
  }
  int64_t sb = bid - p.n_chunk_blocks;
  if (p.xcd_remap) sb = xcd_remap_block(sb, (int64_t)gridDim.x - p.n_chunk_blocks);
  const int64_t tile = sb % p.n_tiles, rb = sb / p.n_tiles;
  const int64_t row = rb * (kMmBlock / G) + threadIdx.x / G;
  const int g = threadIdx.x % G;
  const int64_t n0 = tile * (G * VW) + (int64_t)g * VW;
  if constexpr (kWaveStaged<V, I, G, U>) {
    spmm_staged_rows<A, G, VW, U>(p, rb, row, n0);
    return;
  }
  if (row >= p.n_rows) return;  // the whole group
  const int64_t rs = (int64_t)p.ptr[row], re = (int64_t)p.ptr[row + 1];
  if (re - rs > p.short_max) return;  // chunked elsewhere (the whole group)
  A acc[VW];
#pragma unroll
  for (int k = 0; k < VW; ++k) acc[k].init();
  if constexpr (kGroupLoads<V, I, G, U>) {
    using C = typename A::compute_t;
    constexpr int F = kMmFill<G>;
    __shared__ int32_t slot_col[kMmBlock * F];
    __shared__ C slot_val[kMmBlock * F];
    const int first = (threadIdx.x - g) * F;  // the group's F * G slots
    const bool colok = n0 < p.n_dense;
    mm_accumulate_group<VW, U, G>(acc, slot_col + first, slot_val + first, p.idx, p.val, p.B,
                                  p.ldb, n0, colok, rs, re, g);
    if (!colok) return;
  } else {
    if (n0 >= p.n_dense) return;  // lanes past the last column own nothing
    mm_accumulate<VW, U>(acc, p.idx, p.val, p.B, p.ldb, n0, rs, re, 1);
  }
  write_vec<G, VW>(p.C, row * p.ldc + n0, acc, p.alpha, p.beta);
}

template <typename V, typename A, typename I, typename P, int G, int VW>
int launch_spmm_gv(const fs_csr* M, const void* B, int64_t ldb, void* C, int64_t ldc,
                   int64_t n_dense, double alpha, double beta, const fs_plan* pl, hipStream_t st) {
  SpmmParams<V, I, P> p;
  p.ptr = static_cast<const P*>(M->row_ptr);
  p.idx = static_cast<const I*>(M->col_idx);
  p.val = static_cast<const V*>(M->values);
  p.B = static_cast<const V*>(B);
  p.C = static_cast<V*>(C);
  p.ldb = ldb;
  p.ldc = ldc;
  p.n_dense = n_dense;
  p.n_rows = M->n_rows;
  p.n_tiles = (n_dense + G * VW - 1) / (G * VW);
  p.alpha = alpha;
  p.beta = beta;
  p.short_max = pl->short_max;
  p.n_chunks = pl->n_chunks;
  p.n_chunk_blocks = (pl->n_chunks * p.n_tiles + kMmWaves - 1) / kMmWaves;
  p.chunk_nnz = pl->chunk_nnz;
  p.long_rows = pl->long_rows;
  p.long_chunk_ptr = pl->long_chunk_ptr;
  p.chunk_long = pl->chunk_long;
  p.partials = pl->partials;
  p.xcd_remap = pl->xcd_remap;
  const int64_t rows_per_block = kMmBlock / G;
  const int64_t short_blocks = ((M->n_rows + rows_per_block - 1) / rows_per_block) * p.n_tiles;
  const int64_t grid = p.n_chunk_blocks + short_blocks;
  if (grid == 0) return FS_OK;
  if (grid > 0x7fffffffLL)
    return fail(FS_ERR_UNSUPPORTED, "fs_spmm_csr: grid of %lld blocks too large", (long long)grid);
  // entries in flight per lane: 8 on request for fp32 / fp64 with int32
  // columns (short rows then finish in one or two batches), else 4
  bool launched = false;
  if constexpr (kHotSpmm<V, I>) {
    if (pl->unroll >= 8) {
      spmm_csr_kernel<V, A, I, P, G, VW, 8><<<dim3((unsigned)grid), dim3(kMmBlock), 0, st>>>(p);
      launched = true;
    }
  }
  if (!launched)
    spmm_csr_kernel<V, A, I, P, G, VW, 4><<<dim3((unsigned)grid), dim3(kMmBlock), 0, st>>>(p);
  int rc = check_launch("fs_spmm_csr");
  if (rc || pl->n_chunks <= pl->n_long) return rc;  // no row spans several chunks
  const int64_t blocks = (pl->n_long + kMmWaves - 1) / kMmWaves;
  spmm_fixup_kernel<A><<<dim3((unsigned)blocks), dim3(kMmBlock), 0, st>>>(p, pl->n_long);
  return check_launch("fs_spmm_csr");
}

template <typename V, typename A, typename I, typename P, int VW>
int launch_spmm_g(int G, const fs_csr* M, const void* B, int64_t ldb, void* C, int64_t ldc,
                  int64_t n_dense, double alpha, double beta, const fs_plan* pl, hipStream_t st) {
  if constexpr (kHotSpmm<V, I>) {
    switch (G) {
      case 1: return launch_spmm_gv<V, A, I, P, 1, VW>(M, B, ldb, C, ldc, n_dense, alpha, beta, pl, st);
      case 2: return launch_spmm_gv<V, A, I, P, 2, VW>(M, B, ldb, C, ldc, n_dense, alpha, beta, pl, st);
      case 4: return launch_spmm_gv<V, A, I, P, 4, VW>(M, B, ldb, C, ldc, n_dense, alpha, beta, pl, st);
      case 8: return launch_spmm_gv<V, A, I, P, 8, VW>(M, B, ldb, C, ldc, n_dense, alpha, beta, pl, st);
      case 16: return launch_spmm_gv<V, A, I, P, 16, VW>(M, B, ldb, C, ldc, n_dense, alpha, beta, pl, st);
      case 32: return launch_spmm_gv<V, A, I, P, 32, VW>(M, B, ldb, C, ldc, n_dense, alpha, beta, pl, st);
      default: return launch_spmm_gv<V, A, I, P, 64, VW>(M, B, ldb, C, ldc, n_dense, alpha, beta, pl, st);
    }
  } else {
    if (G <= 4) return launch_spmm_gv<V, A, I, P, 4, VW>(M, B, ldb, C, ldc, n_dense, alpha, beta, pl, st);
    if (G <= 16) return launch_spmm_gv<V, A, I, P, 16, VW>(M, B, ldb, C, ldc, n_dense, alpha, beta, pl, st);
    return launch_spmm_gv<V, A, I, P, 64, VW>(M, B, ldb, C, ldc, n_dense, alpha, beta, pl, st);
  }
}

static inline int pow2_at_least(int64_t v) {
  int g = 1;
  while (g < v && g < 64) g <<= 1;
  return g;
}

template <typename V, typename A, typename I, typename P>
int launch_spmm(const fs_csr* M, const void* B, int64_t ldb, void* C, int64_t ldc, int64_t n_dense,
                double alpha, double beta, const fs_plan* pl, int32_t col_tile, hipStream_t st) {
  // Vector width: the widest of {4, 2, 1} (fp64: {2, 1}; others: 1) that
  // divides n_dense and both leading dimensions and keeps B / C aligned.
  int vw = 1;
  if constexpr (kHotSpmm<V, I>) {
    const int maxv = std::is_same<V, float>::value ? 4 : 2;
    for (int v = maxv; v >= 2; v >>= 1) {
      const uintptr_t al = (uintptr_t)(v * sizeof(V));
      if (n_dense % v == 0 && ldb % v == 0 && ldc % v == 0 && ((uintptr_t)B % al) == 0 &&
          ((uintptr_t)C % al) == 0) {
        vw = v;
        break;
      }
    }
  }
  // a column tile wider than B would leave the lanes past n_dense idle
  int64_t tile = col_tile > 0 && col_tile < n_dense ? col_tile : n_dense;
  if (tile % vw != 0) vw = 1;
  // narrow tiles keep at least 8 lanes per row (narrower loads, more lanes),
  // so a row's entries are loaded a whole batch at a time by its group
  while (vw > 1 && tile / vw < 8) vw >>= 1;
  int G = pow2_at_least((tile + vw - 1) / vw);
  if (vw == 4) return launch_spmm_g<V, A, I, P, 4>(G, M, B, ldb, C, ldc, n_dense, alpha, beta, pl, st);
  if (vw == 2) return launch_spmm_g<V, A, I, P, 2>(G, M, B, ldb, C, ldc, n_dense, alpha, beta, pl, st);
  return launch_spmm_g<V, A, I, P, 1>(G, M, B, ldb, C, ldc, n_dense, alpha, beta, pl, st);
}

template <typename V, typename A>
int launch_spmm_idx(const fs_csr* M, const void* B, int64_t ldb, void* C, int64_t ldc, int64_t n_dense,
                    double alpha, double beta, const fs_plan* pl, int32_t col_tile, hipStream_t st) {
  if (M->index_type == FS_I32 && M->ptr_type == FS_I32) return launch_spmm<V, A, int32_t, int32_t>(M, B, ldb, C, ldc, n_dense, alpha, beta, pl, col_tile, st);
  if (M->index_type == FS_I32 && M->ptr_type == FS_I64) return launch_spmm<V, A, int32_t, int64_t>(M, B, ldb, C, ldc, n_dense, alpha, beta, pl, col_tile, st);
  if (M->index_type == FS_I64 && M->ptr_type == FS_I32) return launch_spmm<V, A, int64_t, int32_t>(M, B, ldb, C, ldc, n_dense, alpha, beta, pl, col_tile, st);
  return launch_spmm<V, A, int64_t, int64_t>(M, B, ldb, C, ldc, n_dense, alpha, beta, pl, col_tile, st);
}

}  // namespace fs

using namespace fs;

extern "C" int fs_spmm_csr(const fs_csr* A, const void* B, int64_t ldb, void* C, int64_t ldc,
                           int64_t n_dense, double alpha, double beta, const fs_plan* plan,
                           int32_t col_tile, void* stream) {
  int rc = check_csr(A, "fs_spmm_csr");
  if (rc) return rc;
  if (n_dense < 0) return fail(FS_ERR_VALUE, "fs_spmm_csr: n_dense must be >= 0");
  if (col_tile < 0) return fail(FS_ERR_VALUE, "fs_spmm_csr: col_tile must be >= 0");
  rc = check_plan(plan, A, n_dense, "fs_spmm_csr");
  if (rc) return rc;
  if (A->n_rows == 0 || n_dense == 0) return FS_OK;
  if (!C) return fail(FS_ERR_VALUE, "fs_spmm_csr: C is null");
  if (A->n_cols > 0 && !B) return fail(FS_ERR_VALUE, "fs_spmm_csr: B is null");
  if (ldc < n_dense || (A->n_cols > 0 && ldb < n_dense))
    return fail(FS_ERR_VALUE, "fs_spmm_csr: leading dimensions must be >= n_dense");
  hipStream_t st = as_stream(stream);
  const bool wide = plan->accumulate == FS_ACC_WIDE;
  switch (A->value_type) {
    case FS_F16: return launch_spmm_idx<f16_t, Plain<float>>(A, B, ldb, C, ldc, n_dense, alpha, beta, plan, col_tile, st);
    case FS_BF16: return launch_spmm_idx<bf16_t, Plain<float>>(A, B, ldb, C, ldc, n_dense, alpha, beta, plan, col_tile, st);
    case FS_F32:
      return wide ? launch_spmm_idx<float, Plain<double>>(A, B, ldb, C, ldc, n_dense, alpha, beta, plan, col_tile, st)
                  : launch_spmm_idx<float, Plain<float>>(A, B, ldb, C, ldc, n_dense, alpha, beta, plan, col_tile, st);
    case FS_F64: return launch_spmm_idx<double, Plain<double>>(A, B, ldb, C, ldc, n_dense, alpha, beta, plan, col_tile, st);
    case FS_C32: return launch_spmm_idx<c32_t, Plain<cplx<float>>>(A, B, ldb, C, ldc, n_dense, alpha, beta, plan, col_tile, st);
    case FS_C64: return launch_spmm_idx<c64_t, Plain<cplx<float>>>(A, B, ldb, C, ldc, n_dense, alpha, beta, plan, col_tile, st);
    case FS_C128: return launch_spmm_idx<c128_t, Plain<cplx<double>>>(A, B, ldb, C, ldc, n_dense, alpha, beta, plan, col_tile, st);
    default: return fail(FS_ERR_TYPE, "fs_spmm_csr: unsupported value_type %d", A->value_type);
  }
}
