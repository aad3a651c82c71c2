// sliced.hip -- the combine pass of the column-sliced SpMV.
//
// A matrix whose x does not fit one XCD's L2 is stored column-slice-major:
// the entries of slice k (columns [k*w, (k+1)*w)) come before those of slice
// k + 1, grouped by row inside a slice.  Each (slice, row) pair with entries
// is a "compact row"; an ordinary SpMV over the compact rows (entry tiles,
// block groups pinned to XCDs so each XCD's L2 holds only its own x columns)
// writes one partial z[i] per compact row.  This pass then forms
//   y[r] = alpha * sum_k z[compact row (k, r)] + beta * y[r]
// with the sum taken in slice order in fp64 (complex: two fp64; with
// fp32_sums, compensated fp32), so the result does not depend on scheduling.
//
// No counterpart in the reference (its SpMV reads x in place,
// spmv_csr.py:158-184); the row partition and the error contract are those of
// fs_spmv_csr.
#include "spmv_kernel.hpp"

namespace fs {

constexpr int kCombineRows = 256;  // rows per block (one per thread)

// bounds[k * (n_blocks + 1) + b]: first compact row of slice k whose row is
// in row block b (rows [b * 256, (b + 1) * 256)).  A row appears at most once
// per slice, so a (slice, block) range holds at most 256 compact rows: thread
// t takes the t-th of each.  The loads of kGroup slices are issued together,
// then added to the LDS row sums one slice at a time.
constexpr int kGroup = 8;

template <typename Z, typename V, typename S>
__global__ __launch_bounds__(kCombineRows) void sliced_combine_kernel(
    const Z* __restrict__ z, const int32_t* __restrict__ compact_row,
    const int64_t* __restrict__ bounds, int32_t n_slices, int64_t n_rows, int64_t n_blocks,
    V* __restrict__ y, double alpha, double beta) {
  using C = typename Native<Z>::T;
  __shared__ S acc[kCombineRows];
  const int64_t b = blockIdx.x;
  const int tid = threadIdx.x;
  const int64_t row0 = b * kCombineRows;
  acc[tid] = zero<S>();
  for (int k0 = 0; k0 < n_slices; k0 += kGroup) {
    C zv[kGroup];
    int rv[kGroup];
#pragma unroll
    for (int u = 0; u < kGroup; ++u) {
      const int k = k0 + u;
      rv[u] = -1;
      zv[u] = zero<C>();
      if (k < n_slices) {
        const int64_t lo = bounds[(int64_t)k * (n_blocks + 1) + b];
        const int64_t hi = bounds[(int64_t)k * (n_blocks + 1) + b + 1];
        const int64_t i = lo + tid;
        if (i < hi) {
          zv[u] = load_as<C>(z, i);
          const int64_t r = (int64_t)compact_row[i] - row0;
          if (r >= 0 && r < kCombineRows) rv[u] = (int)r;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kGroup; ++u) {
      __syncthreads();  // slice order: rows are distinct inside a slice
      if (rv[u] >= 0) acc[rv[u]] = add(acc[rv[u]], to_sum<S>(zv[u]));
    }
  }
  __syncthreads();
  const int64_t row = row0 + tid;
  if (row < n_rows) write_out(y, row, final_value(acc[tid]), alpha, beta);
}

// The same sums from the row masks: thread t owns row 256 b + t and, slice
// by slice, finds its partial (if any) at bounds[k][b] + the number of rows
// of the block before it that have one in slice k.  The mask words of kGroup
// slices, their prefix bit counts and the bounds are staged in LDS; each
// thread then issues its kGroup partial loads together and adds them in
// slice order in registers: no row ids are read and no barriers separate
// the slices.
constexpr int kMaskWords = kCombineRows / 32;

// kRB consecutive 256-row blocks per workgroup (thread t owns row t of each):
// the blocks' mask words and bounds are staged together and each thread has
// kRB * kGroup partial loads in flight.
template <typename Z, typename V, typename S, int kRB>
__global__ __launch_bounds__(kCombineRows) void sliced_combine_mask_kernel(
    const Z* __restrict__ z, const uint32_t* __restrict__ row_mask,
    const int64_t* __restrict__ bounds, int32_t n_slices, int64_t n_rows, int64_t n_blocks,
    int64_t n_compact, V* __restrict__ y, double alpha, double beta) {
  using C = typename Native<Z>::T;
  static_assert(kRB * kGroup * kMaskWords <= 128 && kRB * kGroup <= 128, "staging threads");
  __shared__ uint32_t words[kRB][kGroup][kMaskWords];
  __shared__ int32_t before[kRB][kGroup][kMaskWords];  // set bits in the words before j
  __shared__ int64_t base[kRB][kGroup];
  const int64_t b0 = (int64_t)blockIdx.x * kRB;
  const int tid = threadIdx.x;
  S acc[kRB];
#pragma unroll
  for (int r = 0; r < kRB; ++r) acc[r] = zero<S>();
  if (n_compact == 0) {  // uniform: no partials to read
#pragma unroll
    for (int r = 0; r < kRB; ++r) {
      const int64_t row = (b0 + r) * kCombineRows + tid;
      if (row < n_rows) write_out(y, row, final_value(acc[r]), alpha, beta);
    }
    return;
  }
  const int w = tid / 32, bit = tid % 32;
  const uint32_t below = bit == 0 ? 0u : (0xffffffffu >> (32 - bit));
  for (int k0 = 0; k0 < n_slices; k0 += kGroup) {
    if (k0 > 0) __syncthreads();  // the previous group's words are consumed
    if (tid < kRB * kGroup * kMaskWords) {
      const int r = tid / (kGroup * kMaskWords), u = (tid / kMaskWords) % kGroup;
      const int j = tid % kMaskWords;
      const int64_t b = b0 + r;
      words[r][u][j] = k0 + u < n_slices && b < n_blocks
                           ? row_mask[((int64_t)(k0 + u) * n_blocks + b) * kMaskWords + j]
                           : 0u;
    } else if (tid >= 128 && tid < 128 + kRB * kGroup) {
      const int r = (tid - 128) / kGroup, u = (tid - 128) % kGroup;
      const int64_t b = b0 + r;
      base[r][u] = k0 + u < n_slices && b < n_blocks ? bounds[(int64_t)(k0 + u) * (n_blocks + 1) + b]
                                                     : 0;
    }
    __syncthreads();
    if (tid < kRB * kGroup) {
      const int r = tid / kGroup, u = tid % kGroup;
      int c = 0;
#pragma unroll
      for (int j = 0; j < kMaskWords; ++j) {
        before[r][u][j] = c;
        c += __popc(words[r][u][j]);
      }
    }
    __syncthreads();
    C zv[kRB][kGroup];
    unsigned present[kRB];  // slices of this group holding a partial of the row
#pragma unroll
    for (int r = 0; r < kRB; ++r) {
      present[r] = 0;
#pragma unroll
      for (int u = 0; u < kGroup; ++u) {
        const uint32_t mine = words[r][u][w];
        const int64_t i = base[r][u] + before[r][u][w] + __popc(mine & below);
        const bool has = ((mine >> bit) & 1u) && i < n_compact;  // a bad mask reads nothing
        // every lane loads (absent rows read partial 0) so that the loads
        // stay unconditional and are all in flight before the first add; a
        // load under a per-lane branch is waited for inside the branch
        zv[r][u] = load_as<C>(z, has ? i : 0);
        present[r] |= (unsigned)has << u;
      }
    }
#pragma unroll
    for (int r = 0; r < kRB; ++r)
#pragma unroll
      for (int u = 0; u < kGroup; ++u)
        if ((present[r] >> u) & 1u) acc[r] = add(acc[r], to_sum<S>(zv[r][u]));
  }
#pragma unroll
  for (int r = 0; r < kRB; ++r) {
    const int64_t row = (b0 + r) * kCombineRows + tid;
    if (row < n_rows) write_out(y, row, final_value(acc[r]), alpha, beta);
  }
}

// S: the sum type -- fp64 (two fp64 for complex) by default, a compensated
// fp32 pair for fp32 partials with fp32_sums set.
template <typename Z, typename V, typename S = typename TileSum<typename Native<Z>::T>::T>
int launch_combine(const fs_slices* sl, const void* z, void* y, double alpha, double beta,
                   hipStream_t st) {
  const int64_t nb = (sl->n_rows + kCombineRows - 1) / kCombineRows;
  if (nb == 0) return FS_OK;
  if (sl->row_mask) {
    // two row blocks per workgroup where the partials are at most 8 bytes
    constexpr int kRB = sizeof(typename Native<Z>::T) <= 8 ? 2 : 1;
    sliced_combine_mask_kernel<Z, V, S, kRB><<<dim3((unsigned)((nb + kRB - 1) / kRB)),
                                            dim3(kCombineRows), 0, st>>>(
        static_cast<const Z*>(z), sl->row_mask, sl->bounds, sl->n_slices, sl->n_rows, nb,
        sl->n_compact, static_cast<V*>(y), alpha, beta);
  } else {
    sliced_combine_kernel<Z, V, S><<<dim3((unsigned)nb), dim3(kCombineRows), 0, st>>>(
        static_cast<const Z*>(z), sl->compact_row, sl->bounds, sl->n_slices, sl->n_rows, nb,
        static_cast<V*>(y), alpha, beta);
  }
  return check_launch("fs_sliced_combine");
}

}  // namespace fs

using namespace fs;

extern "C" int fs_sliced_combine(const fs_slices* sl, const void* z, void* y, double alpha,
                                 double beta, void* stream) {
  if (!sl) return fail(FS_ERR_VALUE, "fs_sliced_combine: slices descriptor is null");
  if (sl->n_rows < 0 || sl->n_compact < 0 || sl->n_slices <= 0)
    return fail(FS_ERR_VALUE, "fs_sliced_combine: bad sizes");
  if (sl->rows_per_block != kCombineRows)
    return fail(FS_ERR_VALUE, "fs_sliced_combine: rows_per_block must be %d", kCombineRows);
  if (sl->n_rows == 0) return FS_OK;
  if (!y || !sl->bounds || (sl->n_compact > 0 && (!z || (!sl->compact_row && !sl->row_mask))))
    return fail(FS_ERR_VALUE, "fs_sliced_combine: null pointer");
  hipStream_t st = as_stream(stream);
  const bool sums = sl->z_type != sl->value_type;
  if (sl->fp32_sums && (sl->value_type != FS_F32 || sums))
    return fail(FS_ERR_VALUE, "fs_sliced_combine: fp32_sums is for fp32 partials of fp32 values");
  switch (sl->value_type) {
    case FS_F32:
      if (!sums && sl->fp32_sums) return launch_combine<float, float, f32c>(sl, z, y, alpha, beta, st);
      if (!sums) return launch_combine<float, float>(sl, z, y, alpha, beta, st);
      if (sl->z_type == FS_F64) return launch_combine<double, float>(sl, z, y, alpha, beta, st);
      break;
    case FS_F64:
      if (!sums) return launch_combine<double, double>(sl, z, y, alpha, beta, st);
      break;
    case FS_C64:
      if (!sums) return launch_combine<c64_t, c64_t>(sl, z, y, alpha, beta, st);
      if (sl->z_type == FS_C128) return launch_combine<c128_t, c64_t>(sl, z, y, alpha, beta, st);
      break;
    case FS_C128:
      if (!sums) return launch_combine<c128_t, c128_t>(sl, z, y, alpha, beta, st);
      break;
    default:
      return fail(FS_ERR_TYPE, "fs_sliced_combine: value_type %d is not sliced", sl->value_type);
  }
  return fail(FS_ERR_TYPE, "fs_sliced_combine: partials of type %d cannot feed value_type %d",
              sl->z_type, sl->value_type);
}
