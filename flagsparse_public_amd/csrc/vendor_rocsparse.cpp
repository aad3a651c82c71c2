// vendor_rocsparse.cpp -- rocSPARSE CSR SpMV / SpMM / SDDMM behind a tiny C ABI, built
// as its own library (libfs_vendor.so).  It is the GPU vendor baseline the
// benchmark harness times beside this library's kernels, in the slot where the
// reference's harness times cuSPARSE through CuPy (benchmarks.py:345-507,
// "cusparse_ms"); nothing on the product path links or calls it.
//
// A context holds the rocSPARSE handle, the descriptors and the analysis
// buffer of one (matrix, x/B, y/C) binding; setup runs the buffer-size and
// preprocess stages once, run() only the compute stage on the given stream.
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>

#include "flagsparse_hip.h"

namespace {

thread_local char g_err[256];

int set_err(const char* what, int code) {
  std::snprintf(g_err, sizeof(g_err), "%s failed (status %d)", what, code);
  return code ? code : -1;
}

#define RS(call)                                                       \
  do {                                                                 \
    rocsparse_status st_ = (call);                                     \
    if (st_ != rocsparse_status_success) return set_err(#call, (int)st_); \
  } while (0)

enum Kind { kSpmv, kSpmm, kSddmm };

struct Ctx {
  rocsparse_handle handle = nullptr;
  rocsparse_spmat_descr mat = nullptr;
  rocsparse_dnvec_descr x = nullptr, y = nullptr;
  rocsparse_dnmat_descr B = nullptr, C = nullptr;  // SDDMM: the dense X and Y^T
  rocsparse_datatype compute = rocsparse_datatype_f32_r;
  int alg = 0;
  Kind kind = kSpmv;
  void* buffer = nullptr;
  size_t buffer_bytes = 0;
};

bool map_value(int vt, rocsparse_datatype* out) {
  switch (vt) {
    case FS_F32: *out = rocsparse_datatype_f32_r; return true;
    case FS_F64: *out = rocsparse_datatype_f64_r; return true;
    case FS_C64: *out = rocsparse_datatype_f32_c; return true;
    case FS_C128: *out = rocsparse_datatype_f64_c; return true;
    default: return false;
  }
}

rocsparse_indextype map_index(int it) {
  return it == FS_I64 ? rocsparse_indextype_i64 : rocsparse_indextype_i32;
}

void destroy(Ctx* c) {
  if (!c) return;
  if (c->x) rocsparse_destroy_dnvec_descr(c->x);
  if (c->y) rocsparse_destroy_dnvec_descr(c->y);
  if (c->B) rocsparse_destroy_dnmat_descr(c->B);
  if (c->C) rocsparse_destroy_dnmat_descr(c->C);
  if (c->mat) rocsparse_destroy_spmat_descr(c->mat);
  if (c->handle) rocsparse_destroy_handle(c->handle);
  if (c->buffer) (void)hipFree(c->buffer);
  delete c;
}

// alpha/beta in the compute type: {re, im} pairs cover the complex types.
struct Scalars {
  alignas(16) unsigned char a[16], b[16];
};

Scalars scalars(rocsparse_datatype t, double alpha, double beta) {
  Scalars s;
  std::memset(&s, 0, sizeof(s));
  if (t == rocsparse_datatype_f32_r || t == rocsparse_datatype_f32_c) {
    float fa = (float)alpha, fb = (float)beta;
    std::memcpy(s.a, &fa, 4);
    std::memcpy(s.b, &fb, 4);
  } else {
    std::memcpy(s.a, &alpha, 8);
    std::memcpy(s.b, &beta, 8);
  }
  return s;
}

int make_matrix(Ctx* c, const fs_csr* A, rocsparse_datatype vt) {
  // rocSPARSE takes non-const pointers; it only reads them for SpMV / SpMM
  RS(rocsparse_create_csr_descr(&c->mat, A->n_rows, A->n_cols, A->nnz,
                                const_cast<void*>(A->row_ptr), const_cast<void*>(A->col_idx),
                                const_cast<void*>(A->values), map_index(A->ptr_type),
                                map_index(A->index_type),
                                rocsparse_index_base_zero, vt));
  return 0;
}

int spmv_stage(Ctx* c, rocsparse_spmv_stage stage, const Scalars& s) {
  size_t bytes = c->buffer_bytes;
#pragma GCC diagnostic push
#pragma GCC diagnostic ignored "-Wdeprecated-declarations"
  RS(rocsparse_spmv(c->handle, rocsparse_operation_none, s.a, c->mat, c->x, s.b, c->y, c->compute,
                    (rocsparse_spmv_alg)c->alg, stage, &bytes, c->buffer));
#pragma GCC diagnostic pop
  if (stage == rocsparse_spmv_stage_buffer_size) c->buffer_bytes = bytes;
  return 0;
}

int spmm_stage(Ctx* c, rocsparse_spmm_stage stage, const Scalars& s) {
  size_t bytes = c->buffer_bytes;
  RS(rocsparse_spmm(c->handle, rocsparse_operation_none, rocsparse_operation_none, s.a, c->mat,
                    c->B, s.b, c->C, c->compute, (rocsparse_spmm_alg)c->alg, stage, &bytes,
                    c->buffer));
  if (stage == rocsparse_spmm_stage_buffer_size) c->buffer_bytes = bytes;
  return 0;
}

// SDDMM stages: the buffer query, the preprocess and the compute call share
// their arguments.
int sddmm_stage(Ctx* c, int stage, const Scalars& s) {
  const rocsparse_operation N = rocsparse_operation_none;
  const rocsparse_sddmm_alg alg = (rocsparse_sddmm_alg)c->alg;
  if (stage == 0) {
    size_t bytes = 0;
    RS(rocsparse_sddmm_buffer_size(c->handle, N, N, s.a, c->B, c->C, s.b, c->mat, c->compute, alg,
                                   &bytes));
    c->buffer_bytes = bytes;
  } else if (stage == 1) {
    RS(rocsparse_sddmm_preprocess(c->handle, N, N, s.a, c->B, c->C, s.b, c->mat, c->compute, alg,
                                  c->buffer));
  } else {
    RS(rocsparse_sddmm(c->handle, N, N, s.a, c->B, c->C, s.b, c->mat, c->compute, alg,
                       c->buffer));
  }
  return 0;
}

int alloc_buffer(Ctx* c) {
  if (c->buffer_bytes == 0) return 0;
  if (hipMalloc(&c->buffer, c->buffer_bytes) != hipSuccess) return set_err("hipMalloc", -1);
  return 0;
}

int spmv_setup(Ctx* c, const fs_csr* A, const void* x, void* y, int alg, void* stream) {
  rocsparse_datatype vt;
  if (!map_value(A->value_type, &vt)) return set_err("value type", -2);
  c->compute = vt;
  c->alg = alg;
  RS(rocsparse_create_handle(&c->handle));
  RS(rocsparse_set_stream(c->handle, (hipStream_t)stream));
  int rc = make_matrix(c, A, vt);
  if (rc) return rc;
  RS(rocsparse_create_dnvec_descr(&c->x, A->n_cols, const_cast<void*>(x), vt));
  RS(rocsparse_create_dnvec_descr(&c->y, A->n_rows, y, vt));
  const Scalars s = scalars(vt, 1.0, 0.0);
  if ((rc = spmv_stage(c, rocsparse_spmv_stage_buffer_size, s))) return rc;
  if ((rc = alloc_buffer(c))) return rc;
  return spmv_stage(c, rocsparse_spmv_stage_preprocess, s);
}

int spmm_setup(Ctx* c, const fs_csr* A, const void* B, int64_t n_dense, void* C, int alg,
               void* stream) {
  rocsparse_datatype vt;
  if (!map_value(A->value_type, &vt)) return set_err("value type", -2);
  c->compute = vt;
  c->alg = alg;
  c->kind = kSpmm;
  RS(rocsparse_create_handle(&c->handle));
  RS(rocsparse_set_stream(c->handle, (hipStream_t)stream));
  int rc = make_matrix(c, A, vt);
  if (rc) return rc;
  RS(rocsparse_create_dnmat_descr(&c->B, A->n_cols, n_dense, n_dense, const_cast<void*>(B), vt,
                                  rocsparse_order_row));
  RS(rocsparse_create_dnmat_descr(&c->C, A->n_rows, n_dense, n_dense, C, vt, rocsparse_order_row));
  const Scalars s = scalars(vt, 1.0, 0.0);
  if ((rc = spmm_stage(c, rocsparse_spmm_stage_buffer_size, s))) return rc;
  if ((rc = alloc_buffer(c))) return rc;
  return spmm_stage(c, rocsparse_spmm_stage_preprocess, s);
}

// out[p] = alpha * <x[row p], y[col p]> + beta * out[p]: X = x [n_rows, k]
// row-major, Y^T = y read as a [k, n_cols] column-major matrix, the sampled
// product written into the pattern's values (which must point at out).
int sddmm_setup(Ctx* c, const fs_csr* P, const void* x, const void* y, int64_t k, int alg,
                void* stream) {
  rocsparse_datatype vt;
  if (!map_value(P->value_type, &vt)) return set_err("value type", -2);
  c->compute = vt;
  c->alg = alg;
  c->kind = kSddmm;
  RS(rocsparse_create_handle(&c->handle));
  RS(rocsparse_set_stream(c->handle, (hipStream_t)stream));
  int rc = make_matrix(c, P, vt);
  if (rc) return rc;
  RS(rocsparse_create_dnmat_descr(&c->B, P->n_rows, k, k, const_cast<void*>(x), vt,
                                  rocsparse_order_row));
  RS(rocsparse_create_dnmat_descr(&c->C, k, P->n_cols, k, const_cast<void*>(y), vt,
                                  rocsparse_order_column));
  const Scalars s = scalars(vt, 1.0, 0.0);
  if ((rc = sddmm_stage(c, 0, s))) return rc;
  if ((rc = alloc_buffer(c))) return rc;
  return sddmm_stage(c, 1, s);
}

}  // namespace

extern "C" {

// Binds y = A x for rocSPARSE algorithm `alg` (rocsparse_spmv_alg: 2 adaptive,
// 3 row split, 7 LRB, 8 nnz split); x and y are device vectors of A's value type.
int fsv_spmv_setup(const fs_csr* A, const void* x, void* y, int alg, void* stream, void** ctx) {
  if (!A || !ctx) return set_err("null argument", -3);
  Ctx* c = new (std::nothrow) Ctx;
  if (!c) return set_err("new", -4);
  const int rc = spmv_setup(c, A, x, y, alg, stream);
  if (rc) { destroy(c); *ctx = nullptr; return rc; }
  *ctx = c;
  return 0;
}

// Binds C = A B for row-major B [n_cols, n_dense] and C [n_rows, n_dense];
// `alg` is a rocsparse_spmm_alg value.
int fsv_spmm_setup(const fs_csr* A, const void* B, int64_t n_dense, void* C, int alg,
                   void* stream, void** ctx) {
  if (!A || !ctx) return set_err("null argument", -3);
  Ctx* c = new (std::nothrow) Ctx;
  if (!c) return set_err("new", -4);
  const int rc = spmm_setup(c, A, B, n_dense, C, alg, stream);
  if (rc) { destroy(c); *ctx = nullptr; return rc; }
  *ctx = c;
  return 0;
}

// Binds the sampled product of the pattern P (P->values = the output) with
// x [n_rows, k] and y [n_cols, k], both row-major; `alg` is a
// rocsparse_sddmm_alg value.
int fsv_sddmm_setup(const fs_csr* P, const void* x, const void* y, int64_t k, int alg,
                    void* stream, void** ctx) {
  if (!P || !ctx) return set_err("null argument", -3);
  Ctx* c = new (std::nothrow) Ctx;
  if (!c) return set_err("new", -4);
  const int rc = sddmm_setup(c, P, x, y, k, alg, stream);
  if (rc) { destroy(c); *ctx = nullptr; return rc; }
  *ctx = c;
  return 0;
}

// Enqueues the compute stage (y = alpha A x + beta y, or the SpMM / SDDMM
// analogue).
int fsv_run(void* ctx, double alpha, double beta) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c) return set_err("null context", -3);
  const Scalars s = scalars(c->compute, alpha, beta);
  switch (c->kind) {
    case kSpmm: return spmm_stage(c, rocsparse_spmm_stage_compute, s);
    case kSddmm: return sddmm_stage(c, 2, s);
    default: return spmv_stage(c, rocsparse_spmv_stage_compute, s);
  }
}

void fsv_destroy(void* ctx) { destroy(static_cast<Ctx*>(ctx)); }

const char* fsv_last_error(void) { return g_err; }

}  // extern "C"
