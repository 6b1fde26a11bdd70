// Shared device/host helpers for the gfx950 DDA kernels: complex arithmetic on
// float2/double2, division by run-time invariant integers, launch checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace dda {

template <typename R> struct CxT;
template <> struct CxT<float> { typedef float2 T; };
template <> struct CxT<double> { typedef double2 T; };

template <typename C> __host__ __device__ inline C cmake(double re, double im) {
  C c;
  c.x = (decltype(c.x))re;
  c.y = (decltype(c.y))im;
  return c;
}
template <typename C> __host__ __device__ inline C cadd(C a, C b) { a.x += b.x; a.y += b.y; return a; }
template <typename C> __host__ __device__ inline C csub(C a, C b) { a.x -= b.x; a.y -= b.y; return a; }
template <typename C> __host__ __device__ inline C cmul(C a, C b) {
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
// a * conj(b)
template <typename C> __host__ __device__ inline C cmulc(C a, C b) {
  C r;
  r.x = a.x * b.x + a.y * b.y;
  r.y = a.y * b.x - a.x * b.y;
  return r;
}
template <typename C> __host__ __device__ inline C cscale(C a, decltype(C().x) s) { a.x *= s; a.y *= s; return a; }
template <typename C> __host__ __device__ inline C cconj(C a) { a.y = -a.y; return a; }
// -i * a
template <typename C> __host__ __device__ inline C cmul_mi(C a) { C r; r.x = a.y; r.y = -a.x; return r; }
// +i * a
template <typename C> __host__ __device__ inline C cmul_pi(C a) { C r; r.x = -a.y; r.y = a.x; return r; }
template <typename C> __host__ __device__ inline C czero() { C r; r.x = 0; r.y = 0; return r; }

__host__ __device__ inline float ffma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__host__ __device__ inline double ffma(double a, double b, double c) { return __builtin_fma(a, b, c); }
// a0 x0 + a1 x1 + a2 x2 as one chain of fused multiply-adds per part (12
// instructions; products and sums written separately take 16)
template <typename C> __host__ __device__ inline C cdot3(C a0, C x0, C a1, C x1, C a2, C x2) {
  C r;
  r.x = ffma(a0.x, x0.x, ffma(-a0.y, x0.y, ffma(a1.x, x1.x, ffma(-a1.y, x1.y, ffma(a2.x, x2.x, -a2.y * x2.y)))));
  r.y = ffma(a0.x, x0.y, ffma(a0.y, x0.x, ffma(a1.x, x1.y, ffma(a1.y, x1.x, ffma(a2.x, x2.y, a2.y * x2.x)))));
  return r;
}

// Division of n < 2^31 by an invariant divisor d >= 1 with one multiply-high:
// q = (umulhi(n, m) + n) >> s, m = floor(2^32 (2^s - d) / d) + 1, s = ceil(log2 d)
// (Granlund & Montgomery, "Division by invariant integers using multiplication", 1994).
struct FastDiv {
  uint32_t d;
  uint32_t m;
  uint32_t s;
  __host__ __device__ inline uint32_t div(uint32_t n) const {
#ifdef __HIP_DEVICE_COMPILE__
    return (__umulhi(n, m) + n) >> s;
#else
    return (uint32_t)((((uint64_t)n * m) >> 32) + n) >> s;
#endif
  }
  __host__ __device__ inline uint32_t mod(uint32_t n) const { return n - div(n) * d; }
};

inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  f.d = d;
  uint32_t s = 0;
  while ((1ull << s) < d) s++;
  f.s = s;
  f.m = (uint32_t)((((1ull << s) - d) << 32) / d + 1);
  if (d == 1) { f.m = 0; f.s = 0; }
  return f;
}

}  // namespace dda
