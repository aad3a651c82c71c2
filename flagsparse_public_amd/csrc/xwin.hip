// xwin.hip -- first pass of the x-window SpMV: every entry's x value gathered
// from an x window staged in LDS (entry point fs_xwin_gather).
//
// A random 4-byte x gather that misses L1 is one L2 request; at ~12 entries
// per row over a 16 MB x, those requests -- not HBM bytes -- bound a one-pass
// SpMV.  Here the columns are cut into windows of W columns (W values of x fit
// one workgroup's LDS) and the entries are kept a second time in window-major
// order as 16-bit in-window columns.  A workgroup stages one window of x in
// LDS, then streams that window's entries: each x value is read from LDS (no
// L2 request) and written to xg at the entry's place in the second pass's
// layout.  The second pass (fs_spmv_csr with fs_plan.x_perm) streams values
// and xg in entry order, so both passes are sequential streams.
//
// Entries are grouped into runs: the entries of one entry tile that fall in
// one window, in CSR order.  Within a run, consecutive entries go to
// consecutive xg slots, so entry k of run r goes to xg[k + run_off[r]].  The
// run of an entry is found without a search: batch b (entries 64b .. 64b+63)
// records the run of its first entry and a 64-bit mask of the entries that
// start a run.
//
// That is the tile-major xg.  Its stores land in runs of ~T / n_windows
// entries, and neighbouring runs of one tile (neighbouring windows) are
// written by different workgroups, so most cache lines of xg are written in
// parts, at different times.  The window-major xg (FS_XG_WINDOW_MAJOR)
// writes entry k to xg[k] -- whole lines, in order -- and leaves the run
// lookup to the second pass's loads, which tolerate partial lines.
#include <atomic>
#include <type_traits>

#include "fs_host.hpp"

namespace fs {

constexpr int kMaxXwDevices = 64;
constexpr int kXwBlock = 1024;   // threads per workgroup (16 wavefronts)
constexpr int kXwRound = 16;     // 64-entry batches per wavefront in flight

template <typename V> struct XwParams {
  const V* x;
  V* xg;
  const uint16_t* wcol;
  const int32_t* batch_run;
  const uint64_t* batch_mask;
  const int32_t* run_off;
  const int32_t* part_window;
  const int64_t* part_entry;
  const int64_t* window_entry;
  int64_t n_cols, nnz, n_runs, n_windows;
  int32_t window;
};

// Stage x[c0 .. c0 + nc) (c0 a multiple of the window, itself a multiple of 4)
// in LDS: kStageLoads 16-byte loads per thread issued before any is stored
// (a 128 KB window is one round of loads; every part's waves wait for it).
constexpr int kStageLoads = 8;

template <typename V>
__device__ __forceinline__ void stage_window(const V* __restrict__ x, int64_t c0, int64_t nc, V* xs) {
  const int tid = threadIdx.x;
  constexpr int kPer = 16 / sizeof(V);  // values per 16-byte load
  using Vec = typename std::conditional<sizeof(V) == 4, float4, double2>::type;
  if (((uintptr_t)(x + c0) & 15) == 0) {
    const int64_t nv = nc / kPer;
    const Vec* src = reinterpret_cast<const Vec*>(x + c0);
    Vec* dst = reinterpret_cast<Vec*>(xs);
    for (int64_t i0 = tid; i0 < nv; i0 += (int64_t)kXwBlock * kStageLoads) {
      Vec v[kStageLoads];
#pragma unroll
      for (int u = 0; u < kStageLoads; ++u) {
        const int64_t i = i0 + (int64_t)u * kXwBlock;
        v[u] = src[i < nv ? i : i0];
      }
#pragma unroll
      for (int u = 0; u < kStageLoads; ++u) {
        const int64_t i = i0 + (int64_t)u * kXwBlock;
        if (i < nv) dst[i] = v[u];
      }
    }
    for (int64_t i = kPer * nv + tid; i < nc; i += kXwBlock) xs[i] = x[c0 + i];
  } else {
    for (int64_t i = tid; i < nc; i += kXwBlock) xs[i] = x[c0 + i];
  }
}

// Workgroup p: the window-major entries [part_entry[p], part_entry[p + 1]),
// starting in window part_window[p] and running through as many windows as
// the range covers; for each window it stages that window of x in LDS, then
// gathers for the range's entries inside the window.  Every index read from
// the layout is clamped into its array (LDS window, run table, xg), so a
// damaged layout gives wrong values, never an out-of-bounds access.
template <typename V>
__global__ __launch_bounds__(kXwBlock) void xwin_gather_kernel(XwParams<V> p) {
  extern __shared__ __align__(16) char smem[];
  V* xs = reinterpret_cast<V*>(smem);
  const int64_t part = blockIdx.x;
  int64_t w = p.part_window[part];
  w = w < 0 ? 0 : (w >= p.n_windows ? p.n_windows - 1 : w);
  int64_t k0 = p.part_entry[part], kend = p.part_entry[part + 1];
  k0 = k0 < 0 ? 0 : k0;
  kend = kend > p.nnz ? p.nnz : kend;
  const int tid = threadIdx.x;
  const int lane = tid % kWave;
  const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);
  const int wmax = p.window - 1;
  const int64_t rmax = p.n_runs - 1;
  const uint64_t below = (2ull << lane) - 2;  // bits 1 .. lane (lane 63: bits 1 .. 63)
  constexpr int kWaves = kXwBlock / kWave;
  for (; k0 < kend && w < p.n_windows; ++w) {
    int64_t k1 = p.window_entry[w + 1];
    k1 = k1 > kend ? kend : k1;
    if (k1 <= k0) continue;  // a window without entries in this range (uniform)
    const int64_t c0 = w * (int64_t)p.window;
    int64_t nc = p.n_cols - c0;
    nc = nc > p.window ? p.window : (nc < 0 ? 0 : nc);
    __syncthreads();  // every read of the previous window is done
    stage_window(p.x, c0, nc, xs);
    __syncthreads();
    const int64_t b_first = k0 / kWave, b_end = (k1 + kWave - 1) / kWave;
    for (int64_t b0 = b_first + (int64_t)wave * kXwRound; b0 < b_end;
         b0 += (int64_t)kWaves * kXwRound) {
      // The round's batch metadata with one vector load (lane i < kXwRound
      // takes batch b0 + i), broadcast per batch with readlane; then every
      // column load.  Indices are clamped into the range, and every load is
      // issued before any is used.
      int64_t bl = b0 + (lane < kXwRound ? lane : 0);
      bl = bl < b_end ? bl : b_end - 1;
      const int32_t brv = p.batch_run[bl];
      const uint64_t bmv = p.batch_mask[bl];
      uint16_t wc[kXwRound];
#pragma unroll
      for (int i = 0; i < kXwRound; ++i) {
        int64_t b = b0 + i;
        b = b < b_end ? b : b_end - 1;
        int64_t k = b * kWave + lane;
        k = k < k0 ? k0 : (k >= k1 ? k1 - 1 : k);
        wc[i] = p.wcol[k];
      }
      int32_t off[kXwRound];
#pragma unroll
      for (int i = 0; i < kXwRound; ++i) {
        const int32_t br = __builtin_amdgcn_readlane(brv, i);
        const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)bmv, i);
        const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(bmv >> 32), i);
        const uint64_t bm = ((uint64_t)hi << 32) | lo;
        int64_t r = (int64_t)br + __builtin_popcountll(bm & below);
        r = r < 0 ? 0 : (r > rmax ? rmax : r);
        off[i] = p.run_off[r];
      }
#pragma unroll
      for (int i = 0; i < kXwRound; ++i) {
        const int64_t k = (b0 + i) * kWave + lane;
        const int c = wc[i] > wmax ? wmax : (int)wc[i];
        const V v = xs[c];
        const int64_t d = k + off[i];
        if (b0 + i < b_end && k >= k0 && k < k1 && d >= 0 && d < p.nnz) p.xg[d] = v;
      }
    }
    k0 = k1;
  }
}

// Window-major xg: entry k's value goes to xg[k], so a part's range is one
// stream.  Each thread takes 16 / sizeof(V) consecutive entries per step --
// one load of their in-window columns, LDS reads, one 16-byte store, so a
// wavefront's stores are 1 KB contiguous -- with kWmSteps steps' loads
// issued before the first is used; the unaligned head and tail of each
// window's range (fewer than 8 entries each) go one entry per thread.
// In-window columns are clamped into the window.
constexpr int kWmSteps = 8;

template <typename V> struct WmCols;  // in-window columns of one step
template <> struct WmCols<float> { using T = uint2; };      // 4 x uint16
template <> struct WmCols<double> { using T = uint32_t; };  // 2 x uint16

template <typename V>
__device__ __forceinline__ void wm_unpack(const typename WmCols<V>::T& w, int (&c)[16 / sizeof(V)]) {
  if constexpr (sizeof(V) == 4) {
    c[0] = (int)(w.x & 0xffffu);
    c[1] = (int)(w.x >> 16);
    c[2] = (int)(w.y & 0xffffu);
    c[3] = (int)(w.y >> 16);
  } else {
    c[0] = (int)(w & 0xffffu);
    c[1] = (int)(w >> 16);
  }
}

template <typename V>
__global__ __launch_bounds__(kXwBlock) void xwin_gather_wm_kernel(XwParams<V> p) {
  extern __shared__ __align__(16) char smem[];
  V* xs = reinterpret_cast<V*>(smem);
  const int64_t part = blockIdx.x;
  int64_t w = p.part_window[part];
  w = w < 0 ? 0 : (w >= p.n_windows ? p.n_windows - 1 : w);
  int64_t k0 = p.part_entry[part], kend = p.part_entry[part + 1];
  k0 = k0 < 0 ? 0 : k0;
  kend = kend > p.nnz ? p.nnz : kend;
  const int tid = threadIdx.x;
  const int wmax = p.window - 1;
  for (; k0 < kend && w < p.n_windows; ++w) {
    int64_t k1 = p.window_entry[w + 1];
    k1 = k1 > kend ? kend : k1;
    if (k1 <= k0) continue;  // a window without entries in this range (uniform)
    const int64_t c0 = w * (int64_t)p.window;
    int64_t nc = p.n_cols - c0;
    nc = nc > p.window ? p.window : (nc < 0 ? 0 : nc);
    __syncthreads();  // every read of the previous window is done
    stage_window(p.x, c0, nc, xs);
    __syncthreads();
    // [kb, ke): the 8-aligned body; [k0, kb) and [ke, k1) one entry per thread
    int64_t kb = (k0 + 7) & ~(int64_t)7, ke = k1 & ~(int64_t)7;
    if (kb > ke) kb = ke = k1;
    if (tid < kb - k0) {
      const int c = p.wcol[k0 + tid];
      p.xg[k0 + tid] = xs[c > wmax ? wmax : c];
    }
    if (tid < k1 - ke) {
      const int c = p.wcol[ke + tid];
      p.xg[ke + tid] = xs[c > wmax ? wmax : c];
    }
    constexpr int kPer = 16 / sizeof(V);
    using Vec = typename std::conditional<sizeof(V) == 4, float4, double2>::type;
    using Cw = typename WmCols<V>::T;
    const int64_t ns = (ke - kb) / kPer;
    const Cw* src = reinterpret_cast<const Cw*>(p.wcol + kb);
    Vec* dst = reinterpret_cast<Vec*>(p.xg + kb);
    for (int64_t g0 = tid; g0 < ns; g0 += (int64_t)kXwBlock * kWmSteps) {
      Cw cw[kWmSteps];
#pragma unroll
      for (int u = 0; u < kWmSteps; ++u) {
        const int64_t g = g0 + (int64_t)u * kXwBlock;
        cw[u] = src[g < ns ? g : g0];
      }
#pragma unroll
      for (int u = 0; u < kWmSteps; ++u) {
        const int64_t g = g0 + (int64_t)u * kXwBlock;
        if (g >= ns) break;
        int c[kPer];
        wm_unpack<V>(cw[u], c);
        Vec out;
        V* o = reinterpret_cast<V*>(&out);
#pragma unroll
        for (int i = 0; i < kPer; ++i) o[i] = xs[c[i] > wmax ? wmax : c[i]];
        dst[g] = out;
      }
    }
    k0 = k1;
  }
}

template <typename V>
int launch_xwin(const fs_xwin* L, const void* x, void* xg, hipStream_t st) {
  XwParams<V> p;
  p.x = static_cast<const V*>(x);
  p.xg = static_cast<V*>(xg);
  p.wcol = L->wcol;
  p.batch_run = L->batch_run;
  p.batch_mask = L->batch_mask;
  p.run_off = L->run_off;
  p.part_window = L->part_window;
  p.part_entry = L->part_entry;
  p.window_entry = L->window_entry;
  p.n_cols = L->n_cols;
  p.nnz = L->nnz;
  p.n_runs = L->n_runs;
  p.window = L->window;
  p.n_windows = (L->n_cols + L->window - 1) / L->window;
  const size_t lds = sizeof(V) * (size_t)L->window;
  if (lds > 65536) {  // windows beyond 64 KB of LDS: raise the kernel's dynamic LDS limit
    // once per device and value type, to the device's largest window (the
    // attribute is an upper bound, so later smaller windows need no call)
    static std::atomic<bool> raised[kMaxXwDevices][2];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxXwDevices)
      return fail(FS_ERR_HIP, "fs_xwin_gather: no usable current device");
    std::atomic<bool>& done = raised[dev][sizeof(V) == 8];
    if (!done.load(std::memory_order_acquire)) {
      int most = 0;
      hipDeviceGetAttribute(&most, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
      if ((size_t)most < lds) most = (int)lds;
      hipError_t e = hipFuncSetAttribute((const void*)xwin_gather_kernel<V>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, most);
      if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)xwin_gather_wm_kernel<V>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, most);
      if (e != hipSuccess)
        return fail(FS_ERR_HIP, "fs_xwin_gather: hipFuncSetAttribute: %s", hipGetErrorString(e));
      done.store(true, std::memory_order_release);
    }
  }
  if (L->order == FS_XG_WINDOW_MAJOR)
    xwin_gather_wm_kernel<V><<<dim3((unsigned)L->n_parts), dim3(kXwBlock), lds, st>>>(p);
  else
    xwin_gather_kernel<V><<<dim3((unsigned)L->n_parts), dim3(kXwBlock), lds, st>>>(p);
  return check_launch("fs_xwin_gather");
}

}  // namespace fs

using namespace fs;

extern "C" int64_t fs_xwin_max_window(int32_t value_type) {
  int64_t item;
  switch (value_type) {
    case FS_F32: item = 4; break;
    case FS_F64: item = 8; break;
    default: return -1;
  }
  int dev = 0, bytes = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  if (hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return -1;
  // the whole window in one workgroup's LDS; in-window columns are 16-bit
  int64_t w = bytes / item;
  if (w > 65532) w = 65532;
  return w - w % 4;
}

extern "C" int fs_xwin_gather(const fs_xwin* L, const void* x, void* xg, void* stream) {
  if (!L) return fail(FS_ERR_VALUE, "fs_xwin_gather: layout is null");
  if (L->value_type != FS_F32 && L->value_type != FS_F64)
    return fail(FS_ERR_TYPE, "fs_xwin_gather: x windows are built for fp32 / fp64 values");
  if (L->nnz < 0 || L->n_cols < 0 || L->n_runs < 0 || L->n_parts < 0)
    return fail(FS_ERR_VALUE, "fs_xwin_gather: negative sizes");
  if (L->nnz > 0x7fffffffLL)
    return fail(FS_ERR_UNSUPPORTED, "fs_xwin_gather: more than 2^31 - 1 entries");
  if (L->window <= 0 || L->window % 4 != 0)
    return fail(FS_ERR_VALUE, "fs_xwin_gather: window must be a positive multiple of 4");
  if (L->order != FS_XG_TILE_MAJOR && L->order != FS_XG_WINDOW_MAJOR)
    return fail(FS_ERR_VALUE, "fs_xwin_gather: order must be FS_XG_TILE_MAJOR or "
                "FS_XG_WINDOW_MAJOR");
  if (L->nnz == 0 || L->n_parts == 0) return FS_OK;
  const bool runs = L->order == FS_XG_TILE_MAJOR;  // only the tile-major xg reads runs
  if ((runs && L->n_runs == 0) || L->n_cols == 0)
    return fail(FS_ERR_VALUE, "fs_xwin_gather: entries without runs or columns");
  if (!x || !xg || !L->wcol || !L->part_window || !L->part_entry || !L->window_entry ||
      (runs && (!L->batch_run || !L->batch_mask || !L->run_off)))
    return fail(FS_ERR_VALUE, "fs_xwin_gather: null array");
  if (L->n_parts > 0x7fffffffLL)
    return fail(FS_ERR_UNSUPPORTED, "fs_xwin_gather: too many parts");
  if (!runs && (((uintptr_t)L->wcol | (uintptr_t)xg) & 15) != 0)
    return fail(FS_ERR_VALUE, "fs_xwin_gather: a window-major xg needs wcol and xg 16-byte "
                "aligned");
  const int64_t max_w = fs_xwin_max_window(L->value_type);
  if (L->window > max_w)
    return fail(FS_ERR_UNSUPPORTED, "fs_xwin_gather: window of %d columns exceeds this "
                "device's %lld", L->window, (long long)max_w);
  hipStream_t st = as_stream(stream);
  return L->value_type == FS_F32 ? launch_xwin<float>(L, x, xg, st)
                                 : launch_xwin<double>(L, x, xg, st);
}
