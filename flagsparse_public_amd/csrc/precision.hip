// precision.hip -- split-precision value storage for fp32 CSR SpMV.
//
// Formats (values only; x and y stay fp32, accumulation is fp32 with
// compensated TwoSum summation in the chunked long-row role):
//   FS_STORE_BF16X2      hi = bf16(v), lo = bf16(v - hi): a = hi + lo
//                        (about 16 significant bits, 4 bytes / entry in two
//                        2-byte planes, so a hi-only pass reads half)
//   FS_STORE_BF16        hi plane only (8 significant bits, 2 bytes / entry)
//   FS_STORE_F16_ROWSCALE  h = fp16(v / s_r), s_r = the power of two at or
//                        above max |v| of row r; a = h * s_r, with s_r applied
//                        once per row after the sum (2 bytes / entry + 4 / row)
//
// No counterpart in the reference (whose values are always stored in the
// call's dtype); the kernels reuse the row-split schedule of spmv_kernel.hpp.
#include <math.h>

#include "spmv_kernel.hpp"

namespace fs {

// Value sources of the storage formats.  get4 (entry tiles with 16-byte
// column / value loads) reads four consecutive 2-byte values of a plane with
// one 8-byte load; it needs j % 4 == 0 and 8-byte aligned planes.
struct SplitBf16Vals {
  const uint16_t* hi;
  const uint16_t* lo;
  template <typename C> __device__ __forceinline__ C get(int64_t j) const {
    return (C)(bf16_to_f32(hi[j]) + bf16_to_f32(lo[j]));
  }
  static constexpr bool kVec = true;
  template <typename C> __device__ __forceinline__ void get4(int64_t j, C (&o)[4]) const {
    uint16_t h[4], l[4];
    load4(hi, j, h);
    load4(lo, j, l);
#pragma unroll
    for (int k = 0; k < 4; ++k) o[k] = (C)(bf16_to_f32(h[k]) + bf16_to_f32(l[k]));
  }
  bool vec_aligned() const { return ((uintptr_t)hi % 8) == 0 && ((uintptr_t)lo % 8) == 0; }
  static constexpr bool kRowScale = false;
  __device__ __forceinline__ float row_scale(int64_t) const { return 1.f; }
};

struct HiBf16Vals {
  const uint16_t* hi;
  template <typename C> __device__ __forceinline__ C get(int64_t j) const { return (C)bf16_to_f32(hi[j]); }
  static constexpr bool kVec = true;
  template <typename C> __device__ __forceinline__ void get4(int64_t j, C (&o)[4]) const {
    uint16_t h[4];
    load4(hi, j, h);
#pragma unroll
    for (int k = 0; k < 4; ++k) o[k] = (C)bf16_to_f32(h[k]);
  }
  bool vec_aligned() const { return ((uintptr_t)hi % 8) == 0; }
  static constexpr bool kRowScale = false;
  __device__ __forceinline__ float row_scale(int64_t) const { return 1.f; }
};

struct ScaledF16Vals {
  const uint16_t* h;
  const float* scale;
  template <typename C> __device__ __forceinline__ C get(int64_t j) const { return (C)f16_to_f32(h[j]); }
  static constexpr bool kVec = true;
  template <typename C> __device__ __forceinline__ void get4(int64_t j, C (&o)[4]) const {
    uint16_t v[4];
    load4(h, j, v);
#pragma unroll
    for (int k = 0; k < 4; ++k) o[k] = (C)f16_to_f32(v[k]);
  }
  bool vec_aligned() const { return ((uintptr_t)h % 8) == 0; }
  static constexpr bool kRowScale = true;
  __device__ __forceinline__ float row_scale(int64_t r) const { return scale[r]; }
};

// ---------------------------------------------------------------- encoders

__global__ void encode_bf16_split(const float* __restrict__ v, uint16_t* __restrict__ hi,
                                  uint16_t* __restrict__ lo, int64_t nnz) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nnz; j += stride) {
    const float a = v[j];
    const uint16_t h = f32_to_bf16(a);
    hi[j] = h;
    if (lo) lo[j] = f32_to_bf16(a - bf16_to_f32(h));  // a - hi is exact in fp32
  }
}

// One wavefront per row: row maximum, power-of-two scale, fp16 encode.
template <typename P>
__global__ __launch_bounds__(256) void encode_f16_rowscale(const P* __restrict__ ptr,
                                                           const float* __restrict__ v,
                                                           uint16_t* __restrict__ h,
                                                           float* __restrict__ scale,
                                                           int64_t n_rows) {
  const int lane = threadIdx.x % kWave;
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x / kWave);
  for (int64_t r = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; r < n_rows;
       r += waves) {
    const int64_t s = (int64_t)ptr[r], e = (int64_t)ptr[r + 1];
    float m = 0.f;
    for (int64_t j = s + lane; j < e; j += kWave) m = fmaxf(m, fabsf(v[j]));
    for (int k = 32; k >= 1; k >>= 1) m = fmaxf(m, __shfl_xor(m, k));
    // smallest power of two >= m (1 for an all-zero row); exact division
    float sc = 1.f;
    if (m > 0.f) {
      int ex;
      const float fr = frexpf(m, &ex);  // m = fr * 2^ex, fr in [0.5, 1)
      sc = ldexpf(1.f, fr == 0.5f ? ex - 1 : ex);
    }
    if (lane == 0) scale[r] = sc;
    const float inv = 1.f / sc;  // power of two: exact
    for (int64_t j = s + lane; j < e; j += kWave) h[j] = f32_to_f16(v[j] * inv);
  }
}

// ---------------------------------------------------------------- launches

// Entry tiles over x values gathered per entry (fs_plan.x_perm).
template <typename VS, typename P>
int launch_split_xg(const fs_csr* M, const VS& vs, const void* xg, void* y, double alpha,
                    double beta, const fs_plan* pl, hipStream_t st) {
  SpmvParams<VS, float, uint16_t, P> p;
  p.ptr = static_cast<const P*>(M->row_ptr);
  p.idx = pl->x_perm;
  p.vals = vs;
  p.x = static_cast<const float*>(xg);
  p.y = static_cast<float*>(y);
  p.alpha = alpha;
  p.beta = beta;
  const int64_t grid = spmv_fill_schedule(p, M, pl, 1);
  if (grid > 0x7fffffffLL)
    return fail(FS_ERR_UNSUPPORTED, "fs_spmv_csr_split: grid of %lld blocks too large", (long long)grid);
  return spmv_launch_entry<Comp, true, true>(p, grid, st, "fs_spmv_csr_split");
}

template <typename VS, typename P, int S>
int launch_split_s(const fs_csr* M, const VS& vs, const void* x, void* y, double alpha,
                   double beta, const fs_plan* pl, hipStream_t st) {
  if (pl->x_perm) return launch_split_xg<VS, P>(M, vs, x, y, alpha, beta, pl, st);
  SpmvParams<VS, float, int32_t, P> p;
  p.ptr = static_cast<const P*>(M->row_ptr);
  p.idx = static_cast<const int32_t*>(M->col_idx);
  p.vals = vs;
  p.x = static_cast<const float*>(x);
  p.y = static_cast<float*>(y);
  p.alpha = alpha;
  p.beta = beta;
  const int64_t grid = spmv_fill_schedule(p, M, pl, S);
  if (grid == 0) return FS_OK;
  if (grid > 0x7fffffffLL)
    return fail(FS_ERR_UNSUPPORTED, "fs_spmv_csr_split: grid of %lld blocks too large", (long long)grid);
  if (pl->tile_nnz > 0 && pl->tile_mode == FS_TILE_ENTRIES)
    return spmv_launch_entry<Comp, true>(p, grid, st, "fs_spmv_csr_split");
  if (pl->tile_nnz > 0) {
    if (tile_lds_bytes<double>(pl->tile_nnz) > 65536)
      return fail(FS_ERR_VALUE, "fs_spmv_csr_split: tile_nnz %lld needs more than 64 KB of LDS",
                  (long long)pl->tile_nnz);
    spmv_launch_tiles<Comp, 4>(p, grid, st);
  } else {
    spmv_csr_kernel<Comp, S, 4><<<dim3((unsigned)grid), dim3(kBlock), 0, st>>>(p);
  }
  const int rc = check_launch("fs_spmv_csr_split");
  return rc ? rc : spmv_launch_fixup<Comp>(p, pl, st, "fs_spmv_csr_split");
}

template <typename VS, typename P>
int launch_split(const fs_csr* M, const VS& vs, const void* x, void* y, double alpha, double beta,
                 const fs_plan* pl, hipStream_t st) {
  switch (pl->group_lanes) {
    case 1: return launch_split_s<VS, P, 1>(M, vs, x, y, alpha, beta, pl, st);
    case 2: return launch_split_s<VS, P, 2>(M, vs, x, y, alpha, beta, pl, st);
    case 4: return launch_split_s<VS, P, 4>(M, vs, x, y, alpha, beta, pl, st);
    case 8: return launch_split_s<VS, P, 8>(M, vs, x, y, alpha, beta, pl, st);
    case 16: return launch_split_s<VS, P, 16>(M, vs, x, y, alpha, beta, pl, st);
    case 32: return launch_split_s<VS, P, 32>(M, vs, x, y, alpha, beta, pl, st);
    case 64: return launch_split_s<VS, P, 64>(M, vs, x, y, alpha, beta, pl, st);
    default:
      return fail(FS_ERR_VALUE, "fs_spmv_csr_split: group_lanes must be 1,2,4,8,16,32 or 64, got %d",
                  pl->group_lanes);
  }
}

template <typename VS>
int launch_split_ptr(const fs_csr* M, const VS& vs, const void* x, void* y, double alpha,
                     double beta, const fs_plan* pl, hipStream_t st) {
  if (M->ptr_type == FS_I32) return launch_split<VS, int32_t>(M, vs, x, y, alpha, beta, pl, st);
  return launch_split<VS, int64_t>(M, vs, x, y, alpha, beta, pl, st);
}

static int check_split_csr(const fs_csr* A, const char* who) {
  // the values live in the split planes: check the descriptor as if A->values were set
  fs_csr probe = *A;
  if (!probe.values) probe.values = probe.col_idx;
  int rc = check_csr(&probe, who);
  if (rc) return rc;
  if (A->value_type != FS_F32) return fail(FS_ERR_TYPE, "%s: split storage is for fp32 matrices", who);
  if (A->index_type != FS_I32) return fail(FS_ERR_TYPE, "%s: split storage needs int32 column indices", who);
  return FS_OK;
}

}  // namespace fs

using namespace fs;

extern "C" int fs_split_values(int32_t format, const fs_csr* A, void* hi, void* lo,
                               float* row_scale, void* stream) {
  int rc = check_csr(A, "fs_split_values");
  if (rc) return rc;
  if (A->value_type != FS_F32) return fail(FS_ERR_TYPE, "fs_split_values: values must be fp32");
  hipStream_t st = as_stream(stream);
  const int64_t nnz = A->nnz;
  if (format == FS_STORE_BF16X2 || format == FS_STORE_BF16) {
    if (nnz == 0) return FS_OK;
    if (!hi || (format == FS_STORE_BF16X2 && !lo))
      return fail(FS_ERR_VALUE, "fs_split_values: output planes are null");
    int64_t blocks = (nnz + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    encode_bf16_split<<<(unsigned)blocks, 256, 0, st>>>(
        static_cast<const float*>(A->values), static_cast<uint16_t*>(hi),
        format == FS_STORE_BF16X2 ? static_cast<uint16_t*>(lo) : nullptr, nnz);
    return check_launch("fs_split_values");
  }
  if (format == FS_STORE_F16_ROWSCALE) {
    if (A->n_rows == 0) return FS_OK;
    if (!row_scale || (nnz > 0 && !hi)) return fail(FS_ERR_VALUE, "fs_split_values: outputs are null");
    int64_t blocks = (A->n_rows + 3) / 4;
    if (blocks > 65536) blocks = 65536;
    if (A->ptr_type == FS_I32)
      encode_f16_rowscale<int32_t><<<(unsigned)blocks, 256, 0, st>>>(
          static_cast<const int32_t*>(A->row_ptr), static_cast<const float*>(A->values),
          static_cast<uint16_t*>(hi), row_scale, A->n_rows);
    else
      encode_f16_rowscale<int64_t><<<(unsigned)blocks, 256, 0, st>>>(
          static_cast<const int64_t*>(A->row_ptr), static_cast<const float*>(A->values),
          static_cast<uint16_t*>(hi), row_scale, A->n_rows);
    return check_launch("fs_split_values");
  }
  return fail(FS_ERR_VALUE, "fs_split_values: unknown storage format %d", format);
}

extern "C" int fs_spmv_csr_split(const fs_csr* A, int32_t format, const void* hi, const void* lo,
                                 const float* row_scale, const void* x, void* y, double alpha,
                                 double beta, const fs_plan* plan, void* stream) {
  int rc = check_split_csr(A, "fs_spmv_csr_split");
  if (rc) return rc;
  rc = check_plan(plan, A, 1, "fs_spmv_csr_split");
  if (rc) return rc;
  if (A->n_rows == 0) return FS_OK;
  if (!y && !plan->row_sums) return fail(FS_ERR_VALUE, "fs_spmv_csr_split: y is null");
  if (A->nnz > 0 && (!x || !hi)) return fail(FS_ERR_VALUE, "fs_spmv_csr_split: x or values are null");
  hipStream_t st = as_stream(stream);
  switch (format) {
    case FS_STORE_BF16X2:
      if (A->nnz > 0 && !lo) return fail(FS_ERR_VALUE, "fs_spmv_csr_split: lo plane is null");
      return launch_split_ptr(A, SplitBf16Vals{static_cast<const uint16_t*>(hi), static_cast<const uint16_t*>(lo)},
                              x, y, alpha, beta, plan, st);
    case FS_STORE_BF16:
      return launch_split_ptr(A, HiBf16Vals{static_cast<const uint16_t*>(hi)}, x, y, alpha, beta, plan, st);
    case FS_STORE_F16_ROWSCALE:
      if (!row_scale) return fail(FS_ERR_VALUE, "fs_spmv_csr_split: row_scale is null");
      return launch_split_ptr(A, ScaledF16Vals{static_cast<const uint16_t*>(hi), row_scale}, x, y,
                              alpha, beta, plan, st);
    default:
      return fail(FS_ERR_VALUE, "fs_spmv_csr_split: unknown storage format %d", format);
  }
}
