// fs_host.hpp -- host-side helpers shared by the C-ABI entry points.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/flagsparse_hip.h"
#include "fs_device.hpp"

namespace fs {

// Records the message of the calling thread's last failure and returns code.
int fail(int code, const char* fmt, ...);

inline int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(FS_ERR_HIP, "%s: kernel launch failed: %s", what, hipGetErrorString(e));
  return FS_OK;
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Accumulator partial element size for a value type and accumulation mode.
int64_t partial_elem_bytes(int32_t value_type, int32_t accumulate);

// Validates the structural fields of a CSR descriptor (types, sizes, pointers).
int check_csr(const fs_csr* A, const char* who);
int check_plan(const fs_plan* P, const fs_csr* A, int64_t n_dense, const char* who);

// Blocks of the XCD-aware remap: consecutive hardware block ids rotate over
// the 8 XCDs; the remap gives XCD k the k-th contiguous slice of logical
// blocks so rows that share columns (banded matrices) share an L2.
__device__ __forceinline__ int64_t xcd_remap_block(int64_t b, int64_t nb) {
  constexpr int64_t kXcd = 8;
  const int64_t per = (nb + kXcd - 1) / kXcd;   // logical blocks per XCD slot
  const int64_t xcd = b % kXcd, k = b / kXcd;
  const int64_t full = nb % kXcd == 0 ? kXcd : nb % kXcd;  // XCDs owning `per` blocks
  // XCDs [0, full) own `per` blocks, the rest own per - 1 (when nb % 8 != 0).
  const int64_t start = xcd < full ? xcd * per : full * per + (xcd - full) * (per - 1);
  return start + k;
}

}  // namespace fs
