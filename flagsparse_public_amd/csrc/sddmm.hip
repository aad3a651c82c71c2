// sddmm.hip -- CSR SDDMM for gfx950: out[p] = alpha * <x[row(p)], y[col(p)]> +
// beta * data[p] over the entries p of a CSR pattern (entry point
// fs_sddmm_csr).
//
// Replaces the reference's Triton kernel (sddmm_csr.py:108-151: a program per
// 128 entries, K in BLOCK_K steps, tl.sum reduction).  Here a group of G
// lanes takes kSdE<G> entries: each lane reads VW consecutive elements of the x
// row and of the y row per step (16-byte loads when the rows are aligned),
// keeps a running sum per entry, and the group reduces with DPP / swizzle
// exchanges.  G is the smallest power of two covering K / VW (at most 64), so
// short K does not leave lanes idle.
#include "fs_host.hpp"

namespace fs {

constexpr int kSdBlock = 256;

template <typename T, int VW> struct VecOf;
template <> struct VecOf<float, 4> { using type = float4; };
template <> struct VecOf<float, 1> { using type = float; };
template <> struct VecOf<double, 2> { using type = double2; };
template <> struct VecOf<double, 1> { using type = double; };

template <typename A, typename T, int VW>
__device__ __forceinline__ void sd_fma(A& acc, const T* xr, const T* yr, int64_t k) {
  using C = typename A::compute_t;
  if constexpr (VW == 1) {
    acc.add_prod((C)xr[k], (C)yr[k]);
  } else {
    using V = typename VecOf<T, VW>::type;
    const V a = *reinterpret_cast<const V*>(xr + k);
    const V b = *reinterpret_cast<const V*>(yr + k);
    const T* pa = reinterpret_cast<const T*>(&a);
    const T* pb = reinterpret_cast<const T*>(&b);
#pragma unroll
    for (int i = 0; i < VW; ++i) acc.add_prod((C)pa[i], (C)pb[i]);
  }
}

// Entries per lane group: the group's kSdE entries (kSdBlock / G apart, so
// consecutive groups still read consecutive row ids / columns) have their row
// and column loads, and then their x / y row loads, in flight together.
// Measured on a 1M x 1M Poisson(16) pattern, fp32: one entry per group
// 418 / 880 / 1969 us at K = 32 / 64 / 128; two 333 / 649 / 1472; four
// 334 / 625 / 1324; eight at K = 128 (32-lane groups) 1304 (K = 8: 288, 290,
// 293 us for one, two, four).
template <int G> constexpr int kSdE = G >= 32 ? 8 : (G >= 16 ? 4 : 2);

template <typename A, typename T, typename O, int G, int VW>
__global__ __launch_bounds__(kSdBlock) void sddmm_kernel(
    const int32_t* __restrict__ row_ids, const int32_t* __restrict__ cols,
    const T* __restrict__ x, const T* __restrict__ y, const T* __restrict__ data,
    O* __restrict__ out, int64_t nnz, int64_t k_dim, double alpha, double beta) {
  using C = typename A::compute_t;
  constexpr int kGroups = kSdBlock / G;
  constexpr int kE = kSdE<G>;
  const int64_t p0 = (int64_t)blockIdx.x * kGroups * kE + threadIdx.x / G;
  const int g = threadIdx.x % G;
  const T* xr[kE];
  const T* yr[kE];
  bool live[kE];  // the whole group agrees
#pragma unroll
  for (int e = 0; e < kE; ++e) {
    const int64_t p = p0 + (int64_t)e * kGroups;
    live[e] = p < nnz;
    // past nnz: the last entry again (nnz >= 1), computed and not written,
    // so that no load sits under a per-lane branch
    const int64_t q = live[e] ? p : nnz - 1;
    xr[e] = x + (int64_t)row_ids[q] * k_dim;
    yr[e] = y + (int64_t)cols[q] * k_dim;
  }
  A acc[kE];
#pragma unroll
  for (int e = 0; e < kE; ++e) acc[e].init();
  for (int64_t k = (int64_t)g * VW; k < k_dim; k += (int64_t)G * VW)
#pragma unroll
    for (int e = 0; e < kE; ++e) sd_fma<A, T, VW>(acc[e], xr[e], yr[e], k);
#pragma unroll
  for (int e = 0; e < kE; ++e) {
    group_reduce<G>(acc[e]);
    if (live[e] && g == 0) {
      const int64_t p = p0 + (int64_t)e * kGroups;
      C v = acc[e].value() * (C)alpha;
      if (data) v += (C)data[p] * (C)beta;
      out[p] = (O)v;
    }
  }
}

template <typename A, typename T, typename O, int G>
void sddmm_launch_g(const int32_t* rid, const int32_t* cols, const T* x, const T* y,
                    const T* data, O* out, int64_t nnz, int64_t k, double alpha, double beta,
                    int vw, hipStream_t st) {
  const int64_t per_block = (int64_t)(kSdBlock / G) * kSdE<G>;
  const int64_t blocks = (nnz + per_block - 1) / per_block;
  const dim3 grid((unsigned)blocks), block(kSdBlock);
  constexpr int kVec = sizeof(T) == 4 ? 4 : 2;
  if (vw == kVec)
    sddmm_kernel<A, T, O, G, kVec><<<grid, block, 0, st>>>(rid, cols, x, y, data, out, nnz, k, alpha, beta);
  else
    sddmm_kernel<A, T, O, G, 1><<<grid, block, 0, st>>>(rid, cols, x, y, data, out, nnz, k, alpha, beta);
}

template <typename A, typename T, typename O>
int sddmm_launch(const fs_csr* P, const int32_t* rid, const void* x, const void* y, int64_t k,
                 void* out, double alpha, double beta, hipStream_t st) {
  const int32_t* cols = static_cast<const int32_t*>(P->col_idx);
  const T* xs = static_cast<const T*>(x);
  const T* ys = static_cast<const T*>(y);
  const T* data = static_cast<const T*>(P->values);
  O* o = static_cast<O*>(out);
  constexpr int kVec = sizeof(T) == 4 ? 4 : 2;
  // vector loads need every row start aligned: k a multiple of the width and
  // 16-byte aligned bases
  const bool vec = k % kVec == 0 && ((uintptr_t)x % 16) == 0 && ((uintptr_t)y % 16) == 0;
  const int vw = vec ? kVec : 1;
  int64_t per = (k + vw - 1) / vw;  // lane steps for one entry
  int G = 1;
  while (G < per && G < 64) G <<= 1;
  const int64_t nnz = P->nnz;
  switch (G) {
    case 1: sddmm_launch_g<A, T, O, 1>(rid, cols, xs, ys, data, o, nnz, k, alpha, beta, vw, st); break;
    case 2: sddmm_launch_g<A, T, O, 2>(rid, cols, xs, ys, data, o, nnz, k, alpha, beta, vw, st); break;
    case 4: sddmm_launch_g<A, T, O, 4>(rid, cols, xs, ys, data, o, nnz, k, alpha, beta, vw, st); break;
    case 8: sddmm_launch_g<A, T, O, 8>(rid, cols, xs, ys, data, o, nnz, k, alpha, beta, vw, st); break;
    case 16: sddmm_launch_g<A, T, O, 16>(rid, cols, xs, ys, data, o, nnz, k, alpha, beta, vw, st); break;
    case 32: sddmm_launch_g<A, T, O, 32>(rid, cols, xs, ys, data, o, nnz, k, alpha, beta, vw, st); break;
    default: sddmm_launch_g<A, T, O, 64>(rid, cols, xs, ys, data, o, nnz, k, alpha, beta, vw, st); break;
  }
  return check_launch("fs_sddmm_csr");
}

}  // namespace fs

using namespace fs;

extern "C" int fs_sddmm_csr(const fs_csr* P, const int32_t* row_ids, const void* x,
                            const void* y, int64_t k_dim, void* out, int32_t out_type,
                            int32_t accumulate, double alpha, double beta, void* stream) {
  if (!P) return fail(FS_ERR_VALUE, "fs_sddmm_csr: pattern descriptor is null");
  if (P->value_type != FS_F32 && P->value_type != FS_F64)
    return fail(FS_ERR_TYPE, "fs_sddmm_csr: x / y must be fp32 or fp64");
  if (P->index_type != FS_I32)
    return fail(FS_ERR_TYPE, "fs_sddmm_csr: column indices must be int32");
  if (P->nnz < 0 || P->n_rows < 0 || P->n_cols < 0 || k_dim < 0)
    return fail(FS_ERR_VALUE, "fs_sddmm_csr: negative size");
  if (out_type != P->value_type && out_type != FS_F64)
    return fail(FS_ERR_TYPE, "fs_sddmm_csr: out must be the value type or fp64");
  if (accumulate != FS_ACC_NATIVE && accumulate != FS_ACC_WIDE)
    return fail(FS_ERR_VALUE, "fs_sddmm_csr: accumulate must be FS_ACC_NATIVE or FS_ACC_WIDE");
  if (P->nnz == 0) return FS_OK;
  if (!row_ids || !P->col_idx || !out || (k_dim > 0 && (!x || !y)))
    return fail(FS_ERR_VALUE, "fs_sddmm_csr: null pointer");
  hipStream_t st = as_stream(stream);
  const bool wide = accumulate == FS_ACC_WIDE;
  if (P->value_type == FS_F64)
    return out_type == FS_F64
               ? sddmm_launch<Plain<double>, double, double>(P, row_ids, x, y, k_dim, out, alpha, beta, st)
               : fail(FS_ERR_TYPE, "fs_sddmm_csr: bad out type");
  if (out_type == FS_F64)
    return wide ? sddmm_launch<Plain<double>, float, double>(P, row_ids, x, y, k_dim, out, alpha, beta, st)
                : sddmm_launch<Plain<float>, float, double>(P, row_ids, x, y, k_dim, out, alpha, beta, st);
  return wide ? sddmm_launch<Plain<double>, float, float>(P, row_ids, x, y, k_dim, out, alpha, beta, st)
              : sddmm_launch<Plain<float>, float, float>(P, row_ids, x, y, k_dim, out, alpha, beta, st);
}
