// spmv.hip -- CSR SpMV for gfx950 on plainly stored values:
// y = alpha * A x + beta * y (entry point fs_spmv_csr).
//
// Replaces the reference's row-per-program kernel (spmv_csr.py:158-184, the
// default path) and its bucketed CSR-vector kernels (spmv_csr.py:225-343,
// 393-440, the use_opt path) with the single wave64 row-split schedule of
// spmv_kernel.hpp.
#include "spmv_kernel.hpp"

namespace fs {

// Hot instantiations (fp32 / fp64 values, int32 columns) get every group
// width and unroll depth; the other dtype / index combinations are built for
// group widths {1, 4, 16, 64} and unroll 4 to bound the library size.
template <typename V, typename I>
constexpr bool kHotSpmv = (std::is_same<V, float>::value || std::is_same<V, double>::value) &&
                          std::is_same<I, int32_t>::value;

template <typename V, typename A, typename I, typename P, int S>
int launch_spmv_s(const fs_csr* M, const void* x, void* y, double alpha, double beta,
                  const fs_plan* pl, hipStream_t st) {
  using VS = DirectVals<V>;
  SpmvParams<VS, V, I, P> p;
  p.ptr = static_cast<const P*>(M->row_ptr);
  p.idx = static_cast<const I*>(M->col_idx);
  p.vals.v = static_cast<const V*>(M->values);
  p.x = static_cast<const V*>(x);
  p.y = static_cast<V*>(y);
  p.alpha = alpha;
  p.beta = beta;
  const int64_t grid = spmv_fill_schedule(p, M, pl, S);
  if (grid == 0) return FS_OK;
  if (grid > 0x7fffffffLL)
    return fail(FS_ERR_UNSUPPORTED, "fs_spmv_csr: grid of %lld blocks too large", (long long)grid);
  const dim3 g((unsigned)grid), b(kBlock);
  if (pl->tile_nnz > 0) {
    using Sum = typename TileSum<typename A::compute_t>::T;
    if (tile_lds_bytes<Sum>(pl->tile_nnz) > 65536)
      return fail(FS_ERR_VALUE, "fs_spmv_csr: tile_nnz %lld needs more than 64 KB of LDS",
                  (long long)pl->tile_nnz);
    if constexpr (kHotSpmv<V, I>) {
      switch (pl->unroll) {
        case 2: spmv_launch_tiles<A, 2>(p, grid, st); break;
        case 8: spmv_launch_tiles<A, 8>(p, grid, st); break;
        default: spmv_launch_tiles<A, 4>(p, grid, st); break;
      }
    } else {
      spmv_launch_tiles<A, 4>(p, grid, st);
    }
    const int rc = check_launch("fs_spmv_csr");
    return rc ? rc : spmv_launch_fixup<A>(p, pl, st, "fs_spmv_csr");
  }
  if constexpr (kHotSpmv<V, I>) {
    switch (pl->unroll) {
      case 2: spmv_csr_kernel<A, S, 2><<<g, b, 0, st>>>(p); break;
      case 8: spmv_csr_kernel<A, S, 8><<<g, b, 0, st>>>(p); break;
      default: spmv_csr_kernel<A, S, 4><<<g, b, 0, st>>>(p); break;
    }
  } else {
    spmv_csr_kernel<A, S, 4><<<g, b, 0, st>>>(p);
  }
  const int rc = check_launch("fs_spmv_csr");
  return rc ? rc : spmv_launch_fixup<A>(p, pl, st, "fs_spmv_csr");
}

// Entry tiles over x values gathered per entry (fs_plan.x_perm, fp32 / fp64):
// the permutation takes the place of the column indices.
template <typename V, typename A, typename P>
int launch_spmv_entries_xg(const fs_csr* M, const void* xg, void* y, double alpha, double beta,
                           const fs_plan* pl, hipStream_t st) {
  using VS = DirectVals<V>;
  SpmvParams<VS, V, uint16_t, P> p;
  p.ptr = static_cast<const P*>(M->row_ptr);
  p.idx = pl->x_perm;
  p.vals.v = static_cast<const V*>(M->values);
  p.x = static_cast<const V*>(xg);
  p.y = static_cast<V*>(y);
  p.alpha = alpha;
  p.beta = beta;
  const int64_t grid = spmv_fill_schedule(p, M, pl, 1);
  if (grid > 0x7fffffffLL)
    return fail(FS_ERR_UNSUPPORTED, "fs_spmv_csr: grid of %lld blocks too large", (long long)grid);
  return spmv_launch_entry<A, true, true>(p, grid, st, "fs_spmv_csr");
}

template <typename V, typename A, typename I, typename P>
int launch_spmv_entries(const fs_csr* M, const void* x, void* y, double alpha, double beta,
                        const fs_plan* pl, hipStream_t st) {
  if constexpr (std::is_same<V, float>::value || std::is_same<V, double>::value) {
    if (pl->x_perm) return launch_spmv_entries_xg<V, A, P>(M, x, y, alpha, beta, pl, st);
  }
  using VS = DirectVals<V>;
  SpmvParams<VS, V, I, P> p;
  p.ptr = static_cast<const P*>(M->row_ptr);
  p.idx = static_cast<const I*>(M->col_idx);
  p.vals.v = static_cast<const V*>(M->values);
  p.x = static_cast<const V*>(x);
  p.y = static_cast<V*>(y);
  p.alpha = alpha;
  p.beta = beta;
  const int64_t grid = spmv_fill_schedule(p, M, pl, 1);
  if (grid > 0x7fffffffLL)
    return fail(FS_ERR_UNSUPPORTED, "fs_spmv_csr: grid of %lld blocks too large", (long long)grid);
  return spmv_launch_entry<A, kHotSpmv<V, I>>(p, grid, st, "fs_spmv_csr");
}

template <typename V, typename A, typename I, typename P>
int launch_spmv(const fs_csr* M, const void* x, void* y, double alpha, double beta,
                const fs_plan* pl, hipStream_t st) {
  if (pl->tile_nnz > 0 && pl->tile_mode == FS_TILE_ENTRIES)
    return launch_spmv_entries<V, A, I, P>(M, x, y, alpha, beta, pl, st);
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
  if (!y && !plan->row_sums) return fail(FS_ERR_VALUE, "fs_spmv_csr: y is null");
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
