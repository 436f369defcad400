// nqs_common.h — shared device / host helpers of libnqs_hip.so (gfx950).
//
// Spin convention on the device: configurations are int32 ±1 arrays at the
// ABI and in sample buffers; chain state in the samplers is a bitmask with
// bit i = 1 <-> s_i = -1 (the basis convention of the reference's ED,
// src/nqs/nqs_lanczos.c:16-20).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#define NQS_HIP_CHECK(expr)                                                    \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess) {                                                    \
      fprintf(stderr, "nqs_hip: %s failed at %s:%d: %s\n", #expr, __FILE__,   \
              __LINE__, hipGetErrorString(_e));                                \
      return -1;                                                               \
    }                                                                          \
  } while (0)

namespace nqs {

constexpr int kWave = 64;
constexpr int kMaxSites = 256;      // bitmask chains: up to 4 x 64-bit words
constexpr int kMaxHidden = 1024;    // register-resident hidden slices in MH

// xorshift64 (<<13, >>7, <<17) and its 53-bit uniform, as the reference's
// sampler and initialiser use (src/nqs/nqs_sampler.c:31-48).
__host__ __device__ inline uint64_t xs_next(uint64_t& s) {
  uint64_t x = s;
  x ^= x << 13;
  x ^= x >> 7;
  x ^= x << 17;
  s = x;
  return x;
}
// x mod n for 0 < n < 2^21, exact: (x_hi mod n) 2^32 + x_lo < 2^53 is
// reduced in double precision and corrected by one step
__host__ __device__ inline int mod_small(uint64_t x, uint32_t n) {
  const uint32_t rh = (uint32_t)(x >> 32) % n;
  const uint64_t t = ((uint64_t)rh << 32) | (uint32_t)x;
  const int64_t q = (int64_t)((double)t / (double)n);
  int64_t r = (int64_t)t - q * (int64_t)n;
  if (r < 0) r += n;
  else if (r >= (int64_t)n) r -= n;
  return (int)r;
}
// The same x mod n with rn = 1/n (rounded) supplied, in double precision
// throughout: every operand and intermediate is an integer below 2^53 (hi,
// rh 2^32 + lo), each quotient estimate trunc(t * rn) is within one of
// floor(t / n) (relative error of the product below 2^-52), t - q n comes out
// of one fused multiply-add exactly, and one correction step makes the
// remainder exact; no integer division or 64-bit integer arithmetic.
__host__ __device__ inline int mod_recip(uint64_t x, uint32_t n, double rn) {
  const double dn = (double)n;
  const double th = (double)(uint32_t)(x >> 32);
  double q = trunc(th * rn);
  double rh = fma(-q, dn, th);
  rh += rh < 0.0 ? dn : 0.0;
  rh -= rh >= dn ? dn : 0.0;
  const double t = rh * 4294967296.0 + (double)(uint32_t)x;
  q = trunc(t * rn);
  double r = fma(-q, dn, t);
  r += r < 0.0 ? dn : 0.0;
  r -= r >= dn ? dn : 0.0;
  return (int)r;
}
__host__ __device__ inline double xs_uniform(uint64_t& s) {
  return (double)(xs_next(s) >> 11) * (1.0 / 9007199254740992.0);
}
__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// Sum over a group of G lanes (G a power of two <= 64) with an xor
// butterfly: every lane of the group ends with the bitwise-identical sum.
template <int G, typename T>
__device__ inline T group_allreduce_sum(T v) {
#pragma unroll
  for (int m = 1; m < G; m <<= 1) v += __shfl_xor(v, m, G);
  return v;
}

// Deterministic block reduction (fixed tree) of one double per thread;
// result valid in thread 0.  `red` holds blockDim.x doubles.
__device__ inline double block_sum(double v, double* red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int s = blockDim.x >> 1; s > 0; s >>= 1) {
    if (t < s) red[t] += red[t + s];
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

template <typename R>
struct cplx {
  R re, im;
};

template <typename R>
__device__ inline cplx<R> cmul(cplx<R> a, cplx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// tanh of a complex number, (sinh 2x + i sin 2y) / (cosh 2x + cos 2y),
// saturated to +-1 for |x| > 20.  Numerator and denominator are scaled by
// 2e, e = exp(-2|x|) <= 1: tanh = (sgn(x)(1 - e^2) + 2i e sin 2y) /
// (1 + e^2 + 2e cos 2y), with 1 - e^2 = -m(2 + m), m = expm1(-2|x|), exact
// for small |x| (one exponential and one sincos instead of cosh, sinh, cos
// and sin).
__device__ inline void ctanh_d(double x, double y, double& tr, double& ti) {
  if (fabs(x) > 20.0) {
    tr = x > 0 ? 1.0 : -1.0;
    ti = 0.0;
    return;
  }
  const double m = expm1(-2.0 * fabs(x)), e = 1.0 + m;
  double s2, c2;
  sincos(2.0 * y, &s2, &c2);
  double den = 1.0 + e * e + 2.0 * e * c2;
  if (fabs(den) < 1e-300) den = 1e-300;
  const double r = 1.0 / den;
  tr = (x >= 0.0 ? -m : m) * (2.0 + m) * r;
  ti = 2.0 * e * s2 * r;
}
__device__ inline void ctanh_f(float x, float y, float& tr, float& ti) {
  if (fabsf(x) > 10.0f) {
    tr = x > 0 ? 1.0f : -1.0f;
    ti = 0.0f;
    return;
  }
  float den = coshf(2.0f * x) + cosf(2.0f * y);
  if (fabsf(den) < 1e-30f) den = 1e-30f;
  tr = sinhf(2.0f * x) / den;
  ti = sinf(2.0f * y) / den;
}
template <typename R>
__device__ inline void ctanh(R x, R y, R& tr, R& ti);
template <>
__device__ inline void ctanh<double>(double x, double y, double& tr, double& ti) {
  ctanh_d(x, y, tr, ti);
}
template <>
__device__ inline void ctanh<float>(float x, float y, float& tr, float& ti) {
  ctanh_f(x, y, tr, ti);
}

// The logistic pair of a hidden input theta = x + iy,
//   p = 1 / (1 + e^{-2 theta}),  q = 1 / (1 + e^{2 theta}),  p + q = 1,
// prices a change d of theta without cancellation:
//   cosh(theta + d) / cosh(theta) = p e^{d} + q e^{-d},
// a sum of two positive terms for real theta (the form cosh d + tanh(theta)
// sinh d loses every digit once tanh rounds to +-1 against a large -d), and
// after the change p' = p e^{d} / w, q' = q e^{-d} / w with w the ratio.
// Only the member of smaller modulus is kept ("small": q when x >= 0, else
// p), from z = e^{-2|x|} e^{-+2iy} (|z| <= 1): small = z / (1 + z); the
// other is 1 - small, which keeps its relative accuracy since its modulus is
// at least 1/2.  The flag tells which one small is (1: small = p).
template <typename R>
__device__ inline void logistic_small(R x, R y, R& sr, R& si, int& is_p) {
  const R E = exp((R)-2 * fabs(x));
  R s2 = 0, c2 = 1;
  if (y != (R)0) {
    if constexpr (sizeof(R) == 8) sincos((double)(2 * y), (double*)&s2, (double*)&c2);
    else sincosf((float)(2 * y), (float*)&s2, (float*)&c2);
  }
  const R sg = x >= (R)0 ? (R)1 : (R)-1;
  R den = (R)1 + (R)2 * E * c2 + E * E;   // |1 + z|^2
  const R floor_ = sizeof(R) == 8 ? (R)1e-300 : (R)1e-30;
  if (den < floor_) den = floor_;
  const R r = (R)1 / den;
  sr = (E * c2 + E * E) * r;
  si = -sg * E * s2 * r;
  is_p = x < (R)0 ? 1 : 0;
}
// p and q from small and its flag
template <typename R>
__device__ inline void logistic_unpack(R sr, R si, int is_p, R& pr, R& pi, R& qr, R& qi) {
  const R lr = (R)1 - sr, li = -si;
  pr = is_p ? sr : lr;
  pi = is_p ? si : li;
  qr = is_p ? lr : sr;
  qi = is_p ? li : si;
}

// log|2 cosh z| and arg(2 cosh z) in the overflow-safe form the reference
// uses (src/nqs/nqs_ansatz.c:309-318), FP64.
__device__ inline void logcosh2_c(double x, double y, double& la, double& ar) {
  double ax = fabs(x), e = exp(-2.0 * ax);
  double re = (1.0 + e) * cos(y);
  double im = (x >= 0.0 ? 1.0 : -1.0) * (1.0 - e) * sin(y);
  double mag = hypot(re, im);
  la = ax + (mag > 0.0 ? log(mag) : -1e300);
  ar = atan2(im, re);
}

}  // namespace nqs
