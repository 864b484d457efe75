// Device-side helpers shared by the kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ga {

// complex value in registers
template <typename T>
struct cplx {
  T re, im;
};

template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) {
  return {a.re + b.re, a.im + b.im};
}
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) {
  return {a.re - b.re, a.im - b.im};
}
template <typename T>
__device__ __forceinline__ cplx<T> cconj(cplx<T> a) {
  return {a.re, -a.im};
}
template <typename T>
__device__ __forceinline__ cplx<T> cmuli(cplx<T> a) {  // a * i
  return {-a.im, a.re};
}

// exp(2 pi i q / d) for 0 <= q < d, accurate to the type's rounding
__device__ __forceinline__ cplx<double> expi_frac(int64_t q, int64_t d, double) {
  double s, c;
  sincospi(2.0 * double(q) / double(d), &s, &c);
  return {c, s};
}
__device__ __forceinline__ cplx<float> expi_frac(int64_t q, int64_t d, float) {
  // reduce in double so the float argument stays exact to rounding
  double s, c;
  sincospi(2.0 * double(q) / double(d), &s, &c);
  return {float(c), float(s)};
}

// Legendre seed lambda_mm = c_m sin^m(theta) in the scaled form used by the
// recurrence kernels: value = p * 2^(600 k), k <= 0, p in (2^-601 |c_m|, |c_m|].
// sin^m is formed by binary powering with exact exponent bookkeeping.
constexpr int kScaleBits = 600;

__device__ __forceinline__ void legendre_seed(double cm, double sth, int m, double &p,
                                              int &k) {
  int es;
  double b = frexp(sth, &es);  // sth = b 2^es, b in [0.5, 1)
  double rm = 1.0, bm = b;
  int er = 0, eb = 0;
  int mm = m;
  while (mm) {
    if (mm & 1) {
      int t;
      rm = frexp(rm * bm, &t);
      er += eb + t;
    }
    mm >>= 1;
    if (mm) {
      int t;
      bm = frexp(bm * bm, &t);
      eb = 2 * eb + t;
    }
  }
  const long long E = (long long)er + (long long)m * (long long)es;  // <= 0
  const long long negE = -E;
  k = -int(negE / kScaleBits);
  const int rem = int(E + (long long)kScaleBits * (negE / kScaleBits));  // in (-600, 0]
  p = ldexp(cm * rm, rem);
}

// Seeds of the spin-s recurrences at l0 = max(m, s) for one ring:
// +s: K sin^{m+s}(t/2) cos^{|m-s|}(t/2), -s: K' sin^{|m-s|}(t/2) cos^{m+s}(t/2),
// with K = mant 2^exp (signed mantissa) from the plan's seed table.  The powers
// are formed as mantissa and exact binary exponent: b^n = mant 2^e.
__device__ __forceinline__ void scaled_power(double b, int n, double &mant, long long &e) {
  int eb;
  const double bm0 = frexp(b, &eb);
  double rm = 1.0, bm = bm0;
  long long er = 0, ebb = 0;
  int nn = n;
  while (nn) {
    if (nn & 1) {
      int t;
      rm = frexp(rm * bm, &t);
      er += ebb + t;
    }
    nn >>= 1;
    if (nn) {
      int t;
      bm = frexp(bm * bm, &t);
      ebb = 2 * ebb + t;
    }
  }
  mant = rm;
  e = er + (long long)n * eb;
}

// value = mant 2^E (E <= 0 in use) in the scaled form p 2^(600 k), k <= 0
__device__ __forceinline__ void to_scaled(double mant, long long E, double &p, int &k) {
  if (E > 0) {  // cannot exceed O(1) values, but keep the representation valid
    p = ldexp(mant, int(E));
    k = 0;
    return;
  }
  const long long negE = -E;
  k = -int(negE / kScaleBits);
  const int rem = int(E + (long long)kScaleBits * (negE / kScaleBits));
  p = ldexp(mant, rem);
}

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw 2011), counter-based
struct u32x4 {
  uint32_t v[4];
};

__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = uint64_t(0xD2511F53u) * c.v[0];
    const uint64_t p1 = uint64_t(0xCD9E8D57u) * c.v[2];
    const uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
    const uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
    c = u32x4{{hi1 ^ c.v[1] ^ k0, lo1, hi0 ^ c.v[3] ^ k1, lo0}};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// a double in [0, 1) with 53 random bits from two words (27 + 26 bits)
__device__ __forceinline__ double uniform53(uint32_t a, uint32_t b) {
  return (double(a >> 5) * 67108864.0 + double(b >> 6)) * 0x1p-53;
}

// full 64-bit lane masks
__device__ __forceinline__ bool wave_any(bool v) { return __any(v); }
__device__ __forceinline__ bool wave_all(bool v) { return __all(v); }

}  // namespace ga
