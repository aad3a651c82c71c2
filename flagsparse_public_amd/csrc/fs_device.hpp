// fs_device.hpp -- device-side building blocks for the gfx950 CSR kernels.
//
// Value storage types, accumulators (plain, fp64-widened, compensated), and
// wave64 sub-group reductions built on DPP row operations (quad_perm,
// row_half_mirror, row_mirror), ds_swizzle for the 16-lane exchange and
// ds_bpermute for the 32-lane exchange.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fs {

constexpr int kWave = 64;

// ---------------------------------------------------------------- storage

struct bf16_t { uint16_t bits; };
struct f16_t { uint16_t bits; };
struct c32_t { f16_t re, im; };        // complex32: two fp16 halves
struct c64_t { float re, im; };
struct c128_t { double re, im; };

template <typename T> struct cplx { T re, im; };

__device__ __forceinline__ float bf16_to_f32(uint16_t b) {
  return __uint_as_float(static_cast<uint32_t>(b) << 16);
}
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float f16_to_f32(uint16_t b) {
  _Float16 h;
  __builtin_memcpy(&h, &b, 2);
  return static_cast<float>(h);
}
__device__ __forceinline__ uint16_t f32_to_f16(float f) {
  _Float16 h = static_cast<_Float16>(f);
  uint16_t b;
  __builtin_memcpy(&b, &h, 2);
  return b;
}

// Compute type of a storage type (what one product is formed in) when the
// accumulator is "native": fp16/bf16 widen to fp32 (the reference computes
// them in fp32, spmv_csr.py:116-119), complex32 to complex fp32.
template <typename V> struct Native;
template <> struct Native<f16_t> { using T = float; };
template <> struct Native<bf16_t> { using T = float; };
template <> struct Native<float> { using T = float; };
template <> struct Native<double> { using T = double; };
template <> struct Native<c32_t> { using T = cplx<float>; };
template <> struct Native<c64_t> { using T = cplx<float>; };
template <> struct Native<c128_t> { using T = cplx<double>; };

// Load a stored value as compute type C.
template <typename C> __device__ __forceinline__ C load_as(const f16_t* p, int64_t i) { return (C)f16_to_f32(p[i].bits); }
template <typename C> __device__ __forceinline__ C load_as(const bf16_t* p, int64_t i) { return (C)bf16_to_f32(p[i].bits); }
template <typename C> __device__ __forceinline__ C load_as(const float* p, int64_t i) { return (C)p[i]; }
template <typename C> __device__ __forceinline__ C load_as(const double* p, int64_t i) { return (C)p[i]; }
template <typename C> __device__ __forceinline__ C load_as(const c32_t* p, int64_t i) {
  c32_t v = p[i];
  return C{(decltype(C{}.re))f16_to_f32(v.re.bits), (decltype(C{}.re))f16_to_f32(v.im.bits)};
}
template <typename C> __device__ __forceinline__ C load_as(const c64_t* p, int64_t i) {
  c64_t v = p[i];
  return C{(decltype(C{}.re))v.re, (decltype(C{}.re))v.im};
}
template <typename C> __device__ __forceinline__ C load_as(const c128_t* p, int64_t i) {
  c128_t v = p[i];
  return C{(decltype(C{}.re))v.re, (decltype(C{}.re))v.im};
}

// Store compute value into storage type (round once).
__device__ __forceinline__ void store_from(f16_t* p, int64_t i, float v) { p[i].bits = f32_to_f16(v); }
__device__ __forceinline__ void store_from(f16_t* p, int64_t i, double v) { p[i].bits = f32_to_f16((float)v); }
__device__ __forceinline__ void store_from(bf16_t* p, int64_t i, float v) { p[i].bits = f32_to_bf16(v); }
__device__ __forceinline__ void store_from(bf16_t* p, int64_t i, double v) { p[i].bits = f32_to_bf16((float)v); }
__device__ __forceinline__ void store_from(float* p, int64_t i, float v) { p[i] = v; }
__device__ __forceinline__ void store_from(float* p, int64_t i, double v) { p[i] = (float)v; }
__device__ __forceinline__ void store_from(double* p, int64_t i, double v) { p[i] = v; }
template <typename T> __device__ __forceinline__ void store_from(c32_t* p, int64_t i, cplx<T> v) {
  c32_t o; o.re.bits = f32_to_f16((float)v.re); o.im.bits = f32_to_f16((float)v.im); p[i] = o;
}
template <typename T> __device__ __forceinline__ void store_from(c64_t* p, int64_t i, cplx<T> v) {
  p[i] = c64_t{(float)v.re, (float)v.im};
}
template <typename T> __device__ __forceinline__ void store_from(c128_t* p, int64_t i, cplx<T> v) {
  p[i] = c128_t{(double)v.re, (double)v.im};
}

// ------------------------------------------------------------ arithmetic

template <typename T> __device__ __forceinline__ T zero() { return T(0); }
template <> __device__ __forceinline__ cplx<float> zero<cplx<float>>() { return {0.f, 0.f}; }
template <> __device__ __forceinline__ cplx<double> zero<cplx<double>>() { return {0.0, 0.0}; }

__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }
// Complex multiply-add in the order of spmv_csr.py:210-213.
template <typename T>
__device__ __forceinline__ cplx<T> fmadd(cplx<T> a, cplx<T> b, cplx<T> c) {
  return {c.re + (a.re * b.re - a.im * b.im), c.im + (a.re * b.im + a.im * b.re)};
}
__device__ __forceinline__ float add(float a, float b) { return a + b; }
__device__ __forceinline__ double add(double a, double b) { return a + b; }
template <typename T> __device__ __forceinline__ cplx<T> add(cplx<T> a, cplx<T> b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ float scale(float a, double s) { return a * (float)s; }
__device__ __forceinline__ double scale(double a, double s) { return a * s; }
template <typename T> __device__ __forceinline__ cplx<T> scale(cplx<T> a, double s) { return {a.re * (T)s, a.im * (T)s}; }

// ---------------------------------------------------------- accumulators
//
// Plain<T>: running sum in T.  Comp: fp32 running sum with a Neumaier/TwoSum
// correction term (compensated summation).  All accumulators expose
//   add_prod(a, x), merge(other), value(), and lane exchange via xfer().

template <typename T> struct Plain {
  using compute_t = T;
  T s;
  __device__ __forceinline__ void init() { s = zero<T>(); }
  __device__ __forceinline__ void add_prod(T a, T x) { s = fmadd(a, x, s); }
  __device__ __forceinline__ void add_value(T v) { s = add(s, v); }
  __device__ __forceinline__ void merge(const Plain& o) { s = add(s, o.s); }
  __device__ __forceinline__ T value() const { return s; }
};

struct Comp {
  using compute_t = float;
  float s, c;
  __device__ __forceinline__ void init() { s = 0.f; c = 0.f; }
  __device__ __forceinline__ void add_value(float v) {
    float t = s + v;
    float bb = t - s;
    float err = (s - (t - bb)) + (v - bb);  // TwoSum: exact s + v - t
    s = t;
    c += err;
  }
  __device__ __forceinline__ void add_prod(float a, float x) {
    float p = a * x;
    float pe = __builtin_fmaf(a, x, -p);  // exact product error
    add_value(p);
    c += pe;
  }
  __device__ __forceinline__ void merge(const Comp& o) {
    float t = s + o.s;
    float bb = t - s;
    float err = (s - (t - bb)) + (o.s - bb);
    s = t;
    c += o.c + err;
  }
  __device__ __forceinline__ float value() const { return s + c; }
};

// f32c: a compensated fp32 sum held as a value pair (running sum s and the
// accumulated rounding errors c), the sum type of the tile roles when the
// accumulator is Comp: adding two of them is a TwoSum of the sums plus the
// corrections, so a row summed from pieces keeps fp32 arithmetic throughout.
struct f32c {
  float s, c;
};
template <> __device__ __forceinline__ f32c zero<f32c>() { return {0.f, 0.f}; }
__device__ __forceinline__ f32c add(f32c a, f32c b) {
  const float t = a.s + b.s;
  const float bb = t - a.s;
  const float err = (a.s - (t - bb)) + (b.s - bb);
  return {t, a.c + b.c + err};
}
// scaling by a power of two (the fp16 row scales) is exact on both parts
__device__ __forceinline__ f32c scale(f32c a, double s) { return {a.s * (float)s, a.c * (float)s}; }
// the value a finished sum stands for (other sum types are their own value)
__device__ __forceinline__ float final_value(f32c a) { return a.s + a.c; }
template <typename T> __device__ __forceinline__ T final_value(T a) { return a; }

// -------------------------------------------------------- lane exchanges
//
// xmov<K>(v): value held by the partner lane in the K-step of a butterfly-like
// reduction over groups of 2K lanes.  Partners are lane^1, lane^2 (quad_perm),
// lane^7 (row_half_mirror), lane^15 (row_mirror), lane^16 (ds_swizzle) and
// lane^32 (ds_bpermute).  After the steps 1..K/... every lane of a 2K group
// holds the group total, so mirror partners are as good as xor partners.

template <int K> __device__ __forceinline__ int xmov_i32(int v) {
  if constexpr (K == 1) return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
  else if constexpr (K == 2) return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
  else if constexpr (K == 4) return __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false); // row_half_mirror
  else if constexpr (K == 8) return __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false); // row_mirror
  else if constexpr (K == 16) return __builtin_amdgcn_ds_swizzle(v, 0x401F);                   // lane ^ 16
  else {
    static_assert(K == 32, "bad exchange distance");
    int lane = __lane_id();
    return __builtin_amdgcn_ds_bpermute((lane ^ 32) << 2, v);
  }
}

template <int K> __device__ __forceinline__ float xmov(float v) {
  return __int_as_float(xmov_i32<K>(__float_as_int(v)));
}
template <int K> __device__ __forceinline__ double xmov(double v) {
  int64_t b = __double_as_longlong(v);
  int lo = xmov_i32<K>((int)(b & 0xffffffff));
  int hi = xmov_i32<K>((int)(b >> 32));
  return __longlong_as_double(((int64_t)(uint32_t)lo) | ((int64_t)hi << 32));
}
template <int K, typename T> __device__ __forceinline__ cplx<T> xmov(cplx<T> v) {
  return {xmov<K>(v.re), xmov<K>(v.im)};
}
template <int K, typename T> __device__ __forceinline__ Plain<T> xmov(const Plain<T>& a) {
  Plain<T> o;
  o.s = xmov<K>(a.s);
  return o;
}
template <int K> __device__ __forceinline__ f32c xmov(f32c a) { return {xmov<K>(a.s), xmov<K>(a.c)}; }
template <int K> __device__ __forceinline__ Comp xmov(const Comp& a) {
  Comp o;
  o.s = xmov<K>(a.s);
  o.c = xmov<K>(a.c);
  return o;
}

// Reduce an accumulator over aligned groups of S lanes; every lane of the
// group ends with the group total.
template <int S, typename A> __device__ __forceinline__ void group_reduce(A& a) {
  if constexpr (S >= 2) { A o = xmov<1>(a); a.merge(o); }
  if constexpr (S >= 4) { A o = xmov<2>(a); a.merge(o); }
  if constexpr (S >= 8) { A o = xmov<4>(a); a.merge(o); }
  if constexpr (S >= 16) { A o = xmov<8>(a); a.merge(o); }
  if constexpr (S >= 32) { A o = xmov<16>(a); a.merge(o); }
  if constexpr (S >= 64) { A o = xmov<32>(a); a.merge(o); }
}

// Output combine: y = alpha * acc + beta * y_old (y_old not read when beta == 0).
template <typename V, typename T>
__device__ __forceinline__ void write_out(V* y, int64_t i, T acc, double alpha, double beta) {
  T v = (alpha == 1.0) ? acc : scale(acc, alpha);
  if (beta != 0.0) v = add(v, scale(load_as<T>(y, i), beta));
  store_from(y, i, v);
}

__device__ __forceinline__ int64_t ldg_index(const int32_t* p, int64_t i) { return (int64_t)p[i]; }
__device__ __forceinline__ int64_t ldg_index(const int64_t* p, int64_t i) { return p[i]; }

}  // namespace fs
