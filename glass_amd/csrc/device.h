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

// full 64-bit lane masks
__device__ __forceinline__ bool wave_any(bool v) { return __any(v); }
__device__ __forceinline__ bool wave_all(bool v) { return __all(v); }

}  // namespace ga
