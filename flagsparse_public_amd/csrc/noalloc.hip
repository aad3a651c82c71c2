// noalloc.hip -- fs_spmv_csr_noalloc: fs_spmv_csr for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// fs_spmv_csr's kernels keep intermediate sums in the plan's `partials`
// buffer: one partial per long-row chunk (row-group and row-tile schedules)
// or two carries per entry tile.  Here that buffer is a file-scope device
// array of FS_NOALLOC_SCRATCH_BYTES bytes instead, so the caller passes only
// the matrix, the plan's index arrays, x and y.  Every GPU the library runs
// on holds its own copy of the array until the process ends; its address on
// each device is looked up once (hipGetSymbolAddress, with that device
// current) and kept.  Calls on one GPU share the array: calls that overlap on
// that GPU (two streams, or two host threads) race.  Nothing needs clearing
// between calls: every chunk partial and every carry that a launch reads is
// written earlier in the same call (every tile writes its head carry, a tile
// whose last row runs past it writes its tail carry, every chunk writes its
// partial).
#include <mutex>

#include "fs_host.hpp"

namespace fs {

static __device__ __align__(256) unsigned char scratch[FS_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once per device).
static int scratch_address(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(FS_ERR_HIP, "fs_spmv_csr_noalloc: hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(FS_ERR_UNSUPPORTED, "fs_spmv_csr_noalloc: device %d beyond the %d supported", dev,
                kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(FS_ERR_HIP, "fs_spmv_csr_noalloc: hipGetSymbolAddress(scratch) on device %d: %s",
                  dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return FS_OK;
}

// Bytes of `partials` fs_spmv_csr reads and writes for this plan.
static int64_t scratch_need(const fs_csr* A, const fs_plan* P) {
  if (P->tile_nnz > 0 && P->tile_mode == FS_TILE_ENTRIES)
    return fs_tile_carry_bytes(A->value_type, P->n_tiles);
  if (P->n_long > 0) return partial_elem_bytes(A->value_type, P->accumulate) * P->n_chunks;
  return 0;
}

}  // namespace fs

using namespace fs;

extern "C" int fs_spmv_csr_noalloc(const fs_csr* A, const void* x, void* y, double alpha,
                                   double beta, const fs_plan* plan, void* stream) {
  int rc = check_csr(A, "fs_spmv_csr_noalloc");
  if (rc) return rc;
  if (!plan) return fail(FS_ERR_VALUE, "fs_spmv_csr_noalloc: plan is null");
  if (plan->n_tiles < 0 || plan->n_chunks < 0 || plan->n_long < 0)
    return fail(FS_ERR_VALUE, "fs_spmv_csr_noalloc: negative plan counts");
  const int64_t need = scratch_need(A, plan);
  if (need < 0) return fail(FS_ERR_TYPE, "fs_spmv_csr_noalloc: unsupported value_type %d", A->value_type);
  if (need > (int64_t)FS_NOALLOC_SCRATCH_BYTES)
    return fail(FS_ERR_UNSUPPORTED,
                "fs_spmv_csr_noalloc: the plan needs %lld bytes of partials, more than the "
                "%lld-byte scratch (use fs_spmv_csr with a caller-owned buffer)",
                (long long)need, (long long)FS_NOALLOC_SCRATCH_BYTES);
  void* base = nullptr;
  rc = scratch_address(&base);
  if (rc) return rc;
  fs_plan local = *plan;
  local.partials = base;
  local.partials_bytes = FS_NOALLOC_SCRATCH_BYTES;
  return fs_spmv_csr(A, x, y, alpha, beta, &local, stream);
}
