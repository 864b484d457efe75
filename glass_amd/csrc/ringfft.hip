// Per-ring Fourier stage of the HEALPix transforms (FP32 and FP64).
//
// Every HEALPix ring holds n = 4 r pixels (r = i on polar ring i, r = nside on
// the equatorial belt).  A workgroup handles one ring and a PAIR of maps: the
// two real transforms ride in one complex transform (shell 0 in the real part,
// shell 1 in the imaginary part).
//
// Synthesis (coefficients -> pixels):
//   1. fold: G_k = sum_{m = k mod n} F_m e^{i m phi0} + sum_{m = -k mod n, m > 0}
//      conj(F_m e^{i m phi0}); H_k = G(shell 0)_k + i G(shell 1)_k, with the
//      phase m phi0 = 2 pi (m pnum mod pden) / pden reduced exactly.
//   2. x_j = sum_k H_k e^{2 pi i j k / n} split as n = 4 r:
//      x_{j1 + r j2} = sum_{k2} i^{j2 k2} e^{2 pi i j1 k2 / n} D_{k2}[j1],
//      D_{k2} = length-r inverse DFT of H_{4 k1 + k2}.
//   3. length-r DFTs in LDS: radix-2^2 in-place FFT when r is a power of two,
//      otherwise Bluestein with a power-of-two convolution L >= 2r - 1
//      (forward DIF natural -> bit-reversed, pointwise product with the
//      bit-reversed chirp transform, inverse DIT bit-reversed -> natural, so no
//      permutation pass is needed).
//   4. epilogue: pointwise transform (identity / lognormal / squared normal)
//      fused into the pixel store.
// Analysis is the mirror image (forward DFT of the real pixel pair, split by
// Hermitian symmetry, phase removed, weighted).
//
// The values H are held in registers (each thread owns indices j1 = t + i NT),
// LDS holds only one length-L work array.
#include <type_traits>

#include "common.h"
#include "device.h"
#include "kernels.h"

namespace ga {

namespace {

template <typename T>
struct tab_t;
template <>
struct tab_t<float> {
  using type = float2;
};
template <>
struct tab_t<double> {
  using type = double2;
};

template <typename T>
__device__ __forceinline__ cplx<T> ld(const typename tab_t<T>::type *p, int64_t i) {
  auto v = p[i];
  return {v.x, v.y};
}

__device__ __forceinline__ int ilog2(int x) { return 31 - __clz(x); }

// LDS work arrays insert one pad element after every 32 (8-byte complex floats
// span all 64 banks every 32 elements), which removes the bank conflicts of
// the short-span butterflies and of the bit-reversed scatter.
__device__ __forceinline__ int pad(int i) { return i + (i >> 5); }
__host__ __device__ inline int padded_len(int n) { return n + (n >> 5) + 1; }

__device__ __forceinline__ cplx<float> expi_q(int64_t q, int64_t d, float) {
  float s, c;
  sincospif(2.0f * float(q) / float(d), &s, &c);
  return {c, s};
}
__device__ __forceinline__ cplx<double> expi_q(int64_t q, int64_t d, double) {
  double s, c;
  sincospi(2.0 * double(q) / double(d), &s, &c);
  return {c, s};
}

// In-place forward FFT (e^{-2 pi i}), natural order in, bit-reversed out, on P
// arrays X + p ldx at once (one barrier per stage for all of them).
template <typename T, int NT, int P>
__device__ void fft_dif(cplx<T> *X, int ldx, int L, const typename tab_t<T>::type *tw, int maxL) {
  const int t = threadIdx.x;
  const int logL = ilog2(L);
  int h = L >> 1;
  if (logL & 1) {
    const int stride = maxL / (2 * h);
    const int lh = logL - 1;  // h = 2^lh
    for (int jj = t; jj < h * P; jj += NT) {
      const int p = jj >> lh, j = jj & (h - 1);
      cplx<T> *Y = X + p * ldx;
      cplx<T> w = cconj(ld<T>(tw, int64_t(j) * stride));
      cplx<T> u = Y[pad(j)], v = Y[pad(j + h)];
      Y[pad(j)] = cadd(u, v);
      Y[pad(j + h)] = cmul(csub(u, v), w);
    }
    __syncthreads();
    h >>= 1;
  }
  const int L4 = L >> 2, lL4 = logL - 2;
  for (; h >= 2; h >>= 2) {
    const int q = h >> 1, lq = ilog2(q);
    const int stride = maxL / (2 * h);
    for (int gg = t; gg < L4 * P; gg += NT) {
      const int p = gg >> lL4, g = gg & (L4 - 1);
      cplx<T> *Y = X + p * ldx;
      const int blk = g >> lq, pos = g & (q - 1);
      const int b0 = blk * 2 * h + pos;
      cplx<T> w1 = cconj(ld<T>(tw, int64_t(pos) * stride));
      cplx<T> w2 = cmul(w1, w1);
      cplx<T> w3 = {w1.im, -w1.re};  // w1 * (-i)
      cplx<T> a = Y[pad(b0)], b = Y[pad(b0 + q)], c = Y[pad(b0 + h)], d = Y[pad(b0 + h + q)];
      cplx<T> a1 = cadd(a, c), c1 = cmul(csub(a, c), w1);
      cplx<T> b1 = cadd(b, d), d1 = cmul(csub(b, d), w3);
      Y[pad(b0)] = cadd(a1, b1);
      Y[pad(b0 + q)] = cmul(csub(a1, b1), w2);
      Y[pad(b0 + h)] = cadd(c1, d1);
      Y[pad(b0 + h + q)] = cmul(csub(c1, d1), w2);
    }
    __syncthreads();
  }
}

// In-place inverse FFT (e^{+2 pi i}, unscaled), bit-reversed in, natural out,
// on P arrays X + p ldx.
template <typename T, int NT, int P>
__device__ void fft_dit(cplx<T> *X, int ldx, int L, const typename tab_t<T>::type *tw, int maxL) {
  const int t = threadIdx.x;
  const int logL = ilog2(L);
  int h = 1;
  if (logL & 1) {
    const int L2 = L >> 1, lL2 = logL - 1;
    for (int jj = t; jj < L2 * P; jj += NT) {
      const int p = jj >> lL2, j = jj & (L2 - 1);
      cplx<T> *Y = X + p * ldx;
      cplx<T> u = Y[pad(2 * j)], v = Y[pad(2 * j + 1)];
      Y[pad(2 * j)] = cadd(u, v);
      Y[pad(2 * j + 1)] = csub(u, v);
    }
    __syncthreads();
    h = 2;
  }
  const int L4 = L >> 2, lL4 = logL - 2;
  for (; h < L; h <<= 2) {
    const int stride = maxL / (4 * h);
    const int lh = ilog2(h);
    for (int gg = t; gg < L4 * P; gg += NT) {
      const int p = gg >> lL4, g = gg & (L4 - 1);
      cplx<T> *Y = X + p * ldx;
      const int blk = g >> lh, pos = g & (h - 1);
      const int a0 = blk * 4 * h + pos;
      cplx<T> w4 = ld<T>(tw, int64_t(pos) * stride);
      cplx<T> w = cmul(w4, w4);
      cplx<T> w4i = cmuli(w4);
      cplx<T> a = Y[pad(a0)], b = Y[pad(a0 + h)], c = Y[pad(a0 + 2 * h)], d = Y[pad(a0 + 3 * h)];
      cplx<T> tb = cmul(b, w), td = cmul(d, w);
      cplx<T> a1 = cadd(a, tb), b1 = csub(a, tb);
      cplx<T> c1 = cadd(c, td), d1 = csub(c, td);
      cplx<T> tc = cmul(c1, w4), tdd = cmul(d1, w4i);
      Y[pad(a0)] = cadd(a1, tc);
      Y[pad(a0 + 2 * h)] = csub(a1, tc);
      Y[pad(a0 + h)] = cadd(b1, tdd);
      Y[pad(a0 + 3 * h)] = csub(b1, tdd);
    }
    __syncthreads();
  }
}

__device__ __forceinline__ int blu_len(int r) {
  int L = 1;
  while (L < 2 * r - 1) L <<= 1;
  return L;
}

// Inverse length-r DFT of the values H[i][slot][p] (owned j1 = t + i NT) for
// the P arrays p, in place in the registers; X holds P work arrays of stride
// ldx in LDS.  conj_io conjugates input and output (the forward DFT).
template <typename T, int NT, int MAXC, int P>
__device__ void sub_idft(cplx<T> (&H)[MAXC][4][P], int slot, int r, cplx<T> *X, int ldx,
                         int64_t boff, const typename tab_t<T>::type *blu,
                         const typename tab_t<T>::type *tw, int maxL, bool conj_io) {
  const int t = threadIdx.x;
  __syncthreads();  // X free
  if (boff < 0) {   // r is a power of two
    const int bits = ilog2(r);
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      const int j1 = t + i * NT;
      if (j1 < r) {
        const int pos = bits ? int(__brev(unsigned(j1)) >> (32 - bits)) : 0;
#pragma unroll
        for (int p = 0; p < P; ++p) {
          cplx<T> v = H[i][slot][p];
          X[p * ldx + pad(pos)] = conj_io ? cconj(v) : v;
        }
      }
    }
    __syncthreads();
    fft_dit<T, NT, P>(X, ldx, r, tw, maxL);
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      const int j1 = t + i * NT;
      if (j1 < r) {
#pragma unroll
        for (int p = 0; p < P; ++p) {
          cplx<T> v = X[p * ldx + pad(j1)];
          H[i][slot][p] = conj_io ? cconj(v) : v;
        }
      }
    }
  } else {
    const int L = blu_len(r);
    const typename tab_t<T>::type *c = blu + boff;
    const typename tab_t<T>::type *bh = c + r;
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      const int j1 = t + i * NT;
      if (j1 < r) {
        const cplx<T> cj = ld<T>(c, j1);
#pragma unroll
        for (int p = 0; p < P; ++p) {
          cplx<T> v = H[i][slot][p];
          if (conj_io) v = cconj(v);
          X[p * ldx + pad(j1)] = cmul(v, cj);
        }
      }
    }
    for (int jj = t; jj < (L - r) * P; jj += NT) {
      const int p = jj / (L - r), j = r + jj - p * (L - r);
      X[p * ldx + pad(j)] = cplx<T>{T(0), T(0)};
    }
    __syncthreads();
    fft_dif<T, NT, P>(X, ldx, L, tw, maxL);
    const int lL = ilog2(L);
    for (int jj = t; jj < L * P; jj += NT) {
      const int p = jj >> lL, j = jj & (L - 1);
      X[p * ldx + pad(j)] = cmul(X[p * ldx + pad(j)], ld<T>(bh, j));
    }
    __syncthreads();
    fft_dit<T, NT, P>(X, ldx, L, tw, maxL);
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      const int j1 = t + i * NT;
      if (j1 < r) {
        const cplx<T> cj = ld<T>(c, j1);
#pragma unroll
        for (int p = 0; p < P; ++p) {
          cplx<T> v = cmul(X[p * ldx + pad(j1)], cj);
          H[i][slot][p] = conj_io ? cconj(v) : v;
        }
      }
    }
  }
}

// e^x - 1 in FP32 for the lognormal store.  |x| < 1/2 (nearly every value of
// a field whose variance is well below 1): the Taylor series to x^8 / 8!,
// Horner in FMAs (truncation below 1.2e-8 relative; about one ulp in all, no
// transcendental).  Otherwise e^x = 2^y (1 + d), y = x log2(e) rounded to FP32
// and d = (x log2(e) - y) ln 2 from the exact FMA residual of the product and
// the low part of log2(e), minus 1: there |e^x - 1| >= 0.39, so the
// subtraction loses nothing (a few ulp; overflow gives inf, underflow -1).  The
// second form runs only in waves where some lane needs it.
__device__ __forceinline__ float expm1_map(float x) {
  constexpr float c2 = 1.0f / 2, c3 = 1.0f / 6, c4 = 1.0f / 24, c5 = 1.0f / 120,
                  c6 = 1.0f / 720, c7 = 1.0f / 5040, c8 = 1.0f / 40320;
  float p = fmaf(c8, x, c7);
  p = fmaf(p, x, c6);
  p = fmaf(p, x, c5);
  p = fmaf(p, x, c4);
  p = fmaf(p, x, c3);
  p = fmaf(p, x, c2);
  p = fmaf(p, x, 1.0f);
  float r = p * x;
  const bool big = !(fabsf(x) < 0.5f);  // NaN included
  if (wave_any(big)) {
    constexpr float kLog2eHi = 1.44269502162933350f;      // log2(e) rounded to FP32
    constexpr float kLog2eLo = 1.92596299112661746e-08f;  // log2(e) - kLog2eHi
    constexpr float kLn2 = 0.693147180559945309f;
    const float y = x * kLog2eHi;
    const float ylo = fmaf(x, kLog2eLo, fmaf(x, kLog2eHi, -y));
    const float u = __builtin_amdgcn_exp2f(y);
    const float e = fmaf(u, ylo * kLn2, u) - 1.0f;
    if (big) r = e;
  }
  return r;
}
__device__ __forceinline__ float expm1_t(float x) { return expm1_map(x); }
__device__ __forceinline__ double expm1_t(double x) { return expm1(x); }

template <typename T>
__device__ __forceinline__ T apply_epi(T x, int kind, T lam, T par) {
  if (kind == GA_T_LOGNORMAL) {
    T y = expm1_t(x - par * T(0.5));
    return lam == T(1) ? y : lam * y;
  }
  if (kind == GA_T_SQUARED_NORMAL) {
    T d = x - par;
    T y = d * d - T(1);
    return lam == T(1) ? y : lam * y;
  }
  return x;
}

// ref - v for the fused residual of the analysis iterations (ref = the input
// maps at the same pixels), v unchanged without ref
__device__ __forceinline__ float4 resid4(const float *__restrict__ ref, int64_t off, int j1,
                                         float4 v) {
  if (!ref) return v;
  const float4 r = reinterpret_cast<const float4 *>(ref + off)[j1];
  return float4{r.x - v.x, r.y - v.y, r.z - v.z, r.w - v.w};
}

// grid: 8 * ceil(nrings / 8) * npp blocks, XCD-aware: all pairs of one ring on
// one XCD (block b runs on XCD b % 8), consecutive in that XCD's queue.
__device__ __forceinline__ bool decode_ring_block(int nrings, int npp, int &ring, int &pair) {
  const int b = blockIdx.x;
  const int x = b & 7, kq = b >> 3;
  ring = x + 8 * (kq / npp);
  pair = kq - (kq / npp) * npp;
  return ring < nrings;
}

// Ring buffer layout: shells in groups of G (complex), F offset of (ring, m, s)
//   ((ring * ngrp + s / G) * (lmax + 1) + m) * G + s % G,   ngrp = ceil(S / G).
// G = S is the plain [ring][m][S] layout.  The producers (Legendre stages)
// store F_m e^{i m phi0}: the ring phase is already applied.
// A workgroup synthesises a unit of 2P consecutive shells (P complex
// transforms).  Split of the length n = 4 r inverse DFT used here
// (decimation in frequency): with k = k1 + r k2 and j = 4 j1 + j2,
//   x_{4 j1 + j2} = sum_{k1} w_r^{j1 k1} [ w_n^{j2 k1} sum_{k2} i^{j2 k2} H_{k1 + r k2} ],
// so a thread owning k1 = t + i NT folds the bins k1, k1 + r, k1 + 2r, k1 + 3r
// (neighbouring lanes read neighbouring m: coalesced), does the radix-4 step
// and twiddle in registers, then the four length-r DFTs run in LDS and the
// thread stores pixels 4 j1 .. 4 j1 + 3.
// Aliasing fold for the first ring kernel, per shell: bin k of shell q,
//   G_k = sum_{m = k mod n} F_m + sum_{m = -k mod n, m > 0} conj(F_m)   (m <= lmax),
// direct terms summed before conjugate terms, each in increasing m.
// fold_shells_own: one thread, one bin; the loads of four terms are issued
// together at clamped addresses (one round when n > lmax).
template <typename T, int NQ>
__device__ __forceinline__ void fold_shells_own(const cplx<T> *__restrict__ Fu, int G, int nsh,
                                                int k, int n, int lmax, cplx<T> (&g)[NQ]) {
  const int mc0 = k == 0 ? n : n - k;
  if (n > lmax) {
    cplx<T> fd[NQ], fc[NQ];
    const cplx<T> *a = Fu + int64_t(min(k, lmax)) * G;
    const cplx<T> *c = Fu + int64_t(min(mc0, lmax)) * G;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      fd[q] = a[min(q, nsh - 1)];
      fc[q] = c[min(q, nsh - 1)];
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      if (q >= nsh) continue;  // missing shells stay zero
      if (k <= lmax) g[q] = cadd(g[q], fd[q]);
      if (mc0 <= lmax) g[q] = cadd(g[q], cconj(fc[q]));
    }
    return;
  }
  for (int m0 = k; m0 <= lmax; m0 += 4 * n) {
    cplx<T> f[4][NQ];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const cplx<T> *a = Fu + int64_t(min(m0 + u * n, lmax)) * G;
#pragma unroll
      for (int q = 0; q < NQ; ++q) f[u][q] = a[min(q, nsh - 1)];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (m0 + u * n <= lmax)
#pragma unroll
        for (int q = 0; q < NQ; ++q)
          if (q < nsh) g[q] = cadd(g[q], f[u][q]);
  }
  for (int m0 = mc0; m0 <= lmax; m0 += 4 * n) {
    cplx<T> f[4][NQ];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const cplx<T> *a = Fu + int64_t(min(m0 + u * n, lmax)) * G;
#pragma unroll
      for (int q = 0; q < NQ; ++q) f[u][q] = a[min(q, nsh - 1)];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (m0 + u * n <= lmax)
#pragma unroll
        for (int q = 0; q < NQ; ++q)
          if (q < nsh) g[q] = cadd(g[q], cconj(f[u][q]));
  }
}

// fold_shells_coop (small rings, n <= NT / 2): Q = NT / n threads per bin,
// thread (k, q) sums the chunks j = q, q + Q, ... of m (direct m = j n + k,
// conjugate m = (j + 1) n - k); the partial sums meet in LDS and are added in
// q order.  Result: bin k of shell q at X[q n + k].  X must hold 2 NT NQ
// entries; every thread of the workgroup takes part.
template <typename T, int NT, int NQ>
__device__ __forceinline__ void fold_shells_coop(const cplx<T> *__restrict__ Fu, int G, int nsh,
                                                 int n, int lmax, cplx<T> *X) {
  const int t = threadIdx.x;
  const int Q = NT / n, nitems = n * Q;  // nitems <= NT
  const int jmax = lmax / n;
  cplx<T> d[NQ], c[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) d[q] = c[q] = cplx<T>{T(0), T(0)};
  if (t < nitems) {
    const int qq = t / n, k = t - qq * n;
    for (int j0 = qq; j0 <= jmax; j0 += 4 * Q) {
      cplx<T> fd[4][NQ], fc[4][NQ];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int j = j0 + u * Q;
        const cplx<T> *a = Fu + int64_t(min(j * n + k, lmax)) * G;
        const cplx<T> *b = Fu + int64_t(min((j + 1) * n - k, lmax)) * G;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          fd[u][q] = a[min(q, nsh - 1)];
          fc[u][q] = b[min(q, nsh - 1)];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int j = j0 + u * Q;
        if (j * n + k <= lmax)
#pragma unroll
          for (int q = 0; q < NQ; ++q)
            if (q < nsh) d[q] = cadd(d[q], fd[u][q]);
        if ((j + 1) * n - k <= lmax)
#pragma unroll
          for (int q = 0; q < NQ; ++q)
            if (q < nsh) c[q] = cadd(c[q], cconj(fc[u][q]));
      }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      X[(2 * q) * NT + t] = d[q];
      X[(2 * q + 1) * NT + t] = c[q];
    }
  }
  __syncthreads();
  cplx<T> g[NQ];
  if (t < n) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      cplx<T> sd = {T(0), T(0)}, sc = {T(0), T(0)};
      for (int j = 0; j < Q; ++j) sd = cadd(sd, X[(2 * q) * NT + j * n + t]);
      for (int j = 0; j < Q; ++j) sc = cadd(sc, X[(2 * q + 1) * NT + j * n + t]);
      g[q] = cadd(sd, sc);
    }
  }
  __syncthreads();
  if (t < n) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) X[q * n + t] = g[q];
  }
}

template <typename T, int NT, int MAXC, int P>
__global__ __launch_bounds__(NT) void ring_synth_kernel(
    int lmax, int nrings, const int64_t *__restrict__ ring_start,
    const int32_t *__restrict__ ring_n, const int64_t *__restrict__ blu_off,
    const typename tab_t<T>::type *__restrict__ blu, const typename tab_t<T>::type *__restrict__ tw,
    int maxL, int ldx, const cplx<T> *__restrict__ F, int S, int G, int nunit,
    T *__restrict__ maps, int64_t npix, const ga_epilogue *__restrict__ epi) {
  extern __shared__ char smem_raw[];
  cplx<T> *X = reinterpret_cast<cplx<T> *>(smem_raw);
  int ring, unit;
  if (!decode_ring_block(nrings, nunit, ring, unit)) return;
  const int t = threadIdx.x;
  const int sbase = 2 * P * unit;
  const int n = ring_n[ring], r = n >> 2;
  const int64_t boff = blu_off[r];
  const int ngrp = (S + G - 1) / G;
  // pointer of shell sbase at m = 0; the unit's 2P shells are contiguous
  // because G is a multiple of 2P or G = S
  const cplx<T> *Fu = F + (int64_t(ring) * ngrp + sbase / G) * (lmax + 1) * G + sbase % G;
  int nsh = S - sbase;
  if (nsh > 2 * P) nsh = 2 * P;

  // fold of the ring's coefficients into its n bins (see fold_shells_*):
  // small rings cooperatively through LDS, large rings per thread
  const bool coop = n <= NT / 2;
  if (coop) fold_shells_coop<T, NT, 2 * P>(Fu, G, nsh, n, lmax, X);
  __syncthreads();
  cplx<T> H[MAXC][4][P];
#pragma unroll
  for (int i = 0; i < MAXC; ++i) {
    const int k1 = t + i * NT;
    cplx<T> h[4][P];
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
      cplx<T> g[2 * P];
#pragma unroll
      for (int q = 0; q < 2 * P; ++q) g[q] = cplx<T>{T(0), T(0)};
      if (k1 < r) {
        const int k = k1 + r * k2;
        if (coop) {
#pragma unroll
          for (int q = 0; q < 2 * P; ++q) g[q] = X[q * n + k];
        } else {
          fold_shells_own<T, 2 * P>(Fu, G, nsh, k, n, lmax, g);
        }
      }
#pragma unroll
      for (int p = 0; p < P; ++p)
        h[k2][p] = cplx<T>{g[2 * p].re - g[2 * p + 1].im, g[2 * p].im + g[2 * p + 1].re};
    }
    // radix-4 across k2 and twiddle w_n^{j2 k1} (k1 < r, so w2 = w1^2, w3 = w1^3)
    const cplx<T> w1 = expi_q(k1 < r ? k1 : 0, n, T(0));
    const cplx<T> w2 = cmul(w1, w1);
    const cplx<T> w3 = cmul(w2, w1);
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const cplx<T> s02 = cadd(h[0][p], h[2][p]), d02 = csub(h[0][p], h[2][p]);
      const cplx<T> s13 = cadd(h[1][p], h[3][p]), d13 = cmuli(csub(h[1][p], h[3][p]));
      H[i][0][p] = cadd(s02, s13);
      H[i][1][p] = cmul(cadd(d02, d13), w1);
      H[i][2][p] = cmul(csub(s02, s13), w2);
      H[i][3][p] = cmul(csub(d02, d13), w3);
    }
  }
  __syncthreads();  // the bins are read before the sub-transforms reuse X
  // ---- four length-r inverse DFTs per pair
#pragma unroll
  for (int j2 = 0; j2 < 4; ++j2)
    sub_idft<T, NT, MAXC, P>(H, j2, r, X, ldx, boff, blu, tw, maxL, false);

  // ---- epilogue and store: pixel 4 j1 + j2 of each shell (epilogue
  // parameters are workgroup-uniform loads, read where they are used)
  const int64_t st = ring_start[ring];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int q0 = 2 * p, q1 = 2 * p + 1;
    int k0 = GA_T_NORMAL, kk1 = GA_T_NORMAL;
    T l0 = T(1), a0 = T(0), l1 = T(1), a1 = T(0);
    if (epi) {
      if (q0 < nsh) { k0 = epi[sbase + q0].kind; l0 = T(epi[sbase + q0].lamda); a0 = T(epi[sbase + q0].param); }
      if (q1 < nsh) { kk1 = epi[sbase + q1].kind; l1 = T(epi[sbase + q1].lamda); a1 = T(epi[sbase + q1].param); }
    }
    T *out0 = maps + int64_t(sbase + q0) * npix + st;
    T *out1 = maps + int64_t(sbase + q1) * npix + st;
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      const int j1 = t + i * NT;
      if (j1 < r) {
#pragma unroll
        for (int j2 = 0; j2 < 4; ++j2) {
          const int64_t pix = 4 * int64_t(j1) + j2;
          if (q0 < nsh) out0[pix] = apply_epi<T>(H[i][j2][p].re, k0, l0, a0);
          if (q1 < nsh) out1[pix] = apply_epi<T>(H[i][j2][p].im, kk1, l1, a1);
        }
      }
    }
  }
}

// Analysis: maps pair -> F[ring][m][s] = w e^{-i m phi0} DFT(f)_{m mod n}
template <typename T, int NT, int MAXC>
__global__ __launch_bounds__(NT) void ring_anal_kernel(
    int lmax, int nrings, const int64_t *__restrict__ ring_start,
    const int32_t *__restrict__ ring_n, const int32_t *__restrict__ ring_pnum,
    const int32_t *__restrict__ ring_pden, const int64_t *__restrict__ blu_off,
    const typename tab_t<T>::type *__restrict__ blu, const typename tab_t<T>::type *__restrict__ tw,
    int maxL, const T *__restrict__ maps, int S, int npp, int64_t npix, T weight,
    cplx<T> *__restrict__ F) {
  extern __shared__ char smem_raw[];
  cplx<T> *X = reinterpret_cast<cplx<T> *>(smem_raw);
  int ring, pair;
  if (!decode_ring_block(nrings, npp, ring, pair)) return;
  const int t = threadIdx.x;
  const int s0 = 2 * pair, s1 = s0 + 1;
  const bool has1 = s1 < S;
  const int n = ring_n[ring], r = n >> 2;
  const int pnum = ring_pnum[ring], pden = ring_pden[ring];
  const int64_t boff = blu_off[r];
  const int64_t st = ring_start[ring];
  const T *in0 = maps + int64_t(s0) * npix + st;
  const T *in1 = maps + int64_t(s1) * npix + st;

  // x_{4 j1 + j2} = f0 + i f1; thread owns j1
  cplx<T> H[MAXC][4][1];
#pragma unroll
  for (int i = 0; i < MAXC; ++i) {
    const int j1 = t + i * NT;
#pragma unroll
    for (int j2 = 0; j2 < 4; ++j2) {
      cplx<T> v = {T(0), T(0)};
      if (j1 < r) {
        v.re = in0[4 * j1 + j2];
        v.im = has1 ? in1[4 * j1 + j2] : T(0);
      }
      H[i][j2][0] = v;
    }
  }
  // forward length-r DFTs of each decimated sequence
#pragma unroll
  for (int j2 = 0; j2 < 4; ++j2)
    sub_idft<T, NT, MAXC, 1>(H, j2, r, X, 0, boff, blu, tw, maxL, true);
  __syncthreads();
  // X_{k1 + r k2} = sum_{j2} (-i)^{j2 k2} e^{-2 pi i j2 k1 / n} D_{j2}[k1] -> LDS
#pragma unroll
  for (int i = 0; i < MAXC; ++i) {
    const int k1 = t + i * NT;
    if (k1 < r) {
      const cplx<T> y0 = H[i][0][0];
      const cplx<T> y1 = cmul(H[i][1][0], cconj(expi_q(k1, n, T(0))));
      const cplx<T> y2 = cmul(H[i][2][0], cconj(expi_q((2 * k1) % n, n, T(0))));
      const cplx<T> y3 = cmul(H[i][3][0], cconj(expi_q((3 * k1) % n, n, T(0))));
      const cplx<T> s02 = cadd(y0, y2), d02 = csub(y0, y2);
      const cplx<T> s13 = cadd(y1, y3), d13 = cmuli(csub(y3, y1));  // -i (y1 - y3)
      X[pad(k1)] = cadd(s02, s13);
      X[pad(k1 + r)] = cadd(d02, d13);
      X[pad(k1 + 2 * r)] = csub(s02, s13);
      X[pad(k1 + 3 * r)] = csub(d02, d13);
    }
  }
  __syncthreads();
  // split the two real transforms and unfold into m = 0..lmax
  for (int m = t; m <= lmax; m += NT) {
    const int k = m % n;
    const int kk = k ? n - k : 0;
    const cplx<T> a = X[pad(k)], b = cconj(X[pad(kk)]);
    cplx<T> d0 = {T(0.5) * (a.re + b.re), T(0.5) * (a.im + b.im)};
    cplx<T> d1 = {T(0.5) * (a.im - b.im), T(-0.5) * (a.re - b.re)};  // (a - b) / (2i)
    const cplx<T> ph = pnum ? cconj(expi_q(m % pden, pden, T(0))) : cplx<T>{T(1), T(0)};
    cplx<T> *f = F + (int64_t(ring) * (lmax + 1) + m) * S;
    cplx<T> v0 = cmul(d0, ph);
    f[s0] = cplx<T>{weight * v0.re, weight * v0.im};
    if (has1) {
      cplx<T> v1 = cmul(d1, ph);
      f[s1] = cplx<T>{weight * v1.re, weight * v1.im};
    }
  }
}

// ---------------------------------------------------------------------------
// FP32 synthesis, second form: fused radix-2^K passes.
//
// After the radix-4 split (as above), the four length-r inverse DFTs of a
// complex transform run together on four LDS arrays.  Each pass fuses K <= 4
// radix-2 stages: a thread loads the 2^K elements of one group into
// registers, runs the K stages there (the twiddles of a group are one table
// value per stage times constant rotations by 2^(K-s)-th roots of unity),
// and writes them back to the same places, so a 4096-point transform takes 3
// passes and 3 barriers (instead of 6 LDS stages), and a thread never holds
// more than one group.  Power-of-two r: inputs stored bit-reversed, inverse
// decimation in time -> natural order.  Other r (Bluestein): forward
// decimation in frequency (natural -> bit-reversed), product with the
// bit-reversed chirp spectrum, inverse decimation in time.  The two ring
// classes are launched separately (ring lists), each with LDS for its own
// transform length.

// e^{2 pi i k / 16} for k = 0..15: cos and sin
__device__ __forceinline__ float c16(int k) {
  switch (k & 15) {
    case 0: return 1.0f;
    case 1: case 15: return 0.92387953251128674f;
    case 2: case 14: return 0.70710678118654752f;
    case 3: case 13: return 0.38268343236508977f;
    case 4: case 12: return 0.0f;
    case 5: case 11: return -0.38268343236508977f;
    case 6: case 10: return -0.70710678118654752f;
    case 7: case 9: return -0.92387953251128674f;
    default: return -1.0f;
  }
}
__device__ __forceinline__ float s16(int k) { return c16(k - 4); }

// v e^{SIGN 2 pi i k / 16} (k a compile-time constant after unrolling)
template <int SIGN>
__device__ __forceinline__ cplx<float> rot16(cplx<float> v, int k) {
  k &= 15;
  if (k == 0) return v;
  if (k == 4) return SIGN > 0 ? cplx<float>{-v.im, v.re} : cplx<float>{v.im, -v.re};
  if (k == 8) return cplx<float>{-v.re, -v.im};
  if (k == 12) return SIGN > 0 ? cplx<float>{v.im, -v.re} : cplx<float>{-v.im, v.re};
  const float c = c16(k), sn = SIGN * s16(k);
  return cplx<float>{v.re * c - v.im * sn, v.re * sn + v.im * c};
}

__device__ __forceinline__ cplx<float> twid(const float2 *__restrict__ tw, int idx, int sign) {
  const float2 w = tw[idx];
  return cplx<float>{w.x, sign * w.y};
}

// One pass fusing K radix-2 stages with half-spans h0, 2 h0, .., 2^(K-1) h0
// (DIT: stages in increasing span; DIF: decreasing) on A arrays of length N.
// Group = elements base + q h0 (q < 2^K), base = blk 2^K h0 + b, b < h0.
// A group either lies inside one 32-element block or starts on a multiple of
// 32 with h0 dividing 32 or a multiple of it, so its padded slots are
//   pad(base + q h0) = pad(base) + q h0 + (q h0 >> 5),
// offsets that are the same for every thread of the pass.  tw holds
// e^{2 pi i j / 2^ltw}.
template <int K, int NT, bool DIT, int SIGN>
__device__ __forceinline__ void fused_pass(cplx<float> *X, int ldx, int A, int N, int h0,
                                           const float2 *__restrict__ tw, int ltw) {
  constexpr int R = 1 << K;
  const int lh0 = ilog2(h0);
  const int lg = ilog2(N) - K;  // groups per array = N / R
  const int total = A << lg;
  int off[R];
#pragma unroll
  for (int q = 0; q < R; ++q) off[q] = (q << lh0) + ((q << lh0) >> 5);
  for (int gg = threadIdx.x; gg < total; gg += NT) {
    const int a = gg >> lg, g = gg & ((1 << lg) - 1);
    const int b = g & (h0 - 1), blk = g >> lh0;
    const int base = (blk << (K + lh0)) + b;
    cplx<float> *Y = X + a * ldx + pad(base);
    cplx<float> v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = Y[off[q]];
    // ws[s] = e^{SIGN 2 pi i b / (2^(s+1) h0)}: one table value per stage
    cplx<float> ws[K];
#pragma unroll
    for (int s = 0; s < K; ++s) ws[s] = twid(tw, b << (ltw - s - 1 - lh0), SIGN);
#pragma unroll
    for (int ss = 0; ss < K; ++ss) {
      const int s = DIT ? ss : K - 1 - ss;
      const int d = 1 << s;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        if (q & d) continue;
        // twiddle of this pair: ws[s] e^{SIGN 2 pi i (q mod d) / (2 d)}
        const cplx<float> t = rot16<SIGN>(ws[s], (q & (d - 1)) * (16 >> (s + 1)));
        if (DIT) {
          const cplx<float> y = cmul(v[q + d], t);
          const cplx<float> x = v[q];
          v[q] = cadd(x, y);
          v[q + d] = csub(x, y);
        } else {
          const cplx<float> x = v[q], y = v[q + d];
          v[q] = cadd(x, y);
          v[q + d] = cmul(csub(x, y), t);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < R; ++q) Y[off[q]] = v[q];
  }
  __syncthreads();
}

// full power-of-two transform of A arrays: DIT (bit-reversed in, natural out)
// or DIF (natural in, bit-reversed out)
template <int NT, bool DIT, int SIGN>
__device__ __forceinline__ void fused_fft(cplx<float> *X, int ldx, int A, int N,
                                          const float2 *__restrict__ tw, int twlen) {
  const int lN = ilog2(N), ltw = ilog2(twlen);
  if (DIT) {
    for (int l = 0; l < lN;) {
      const int K = lN - l >= 4 ? 4 : lN - l;
      switch (K) {
        case 4: fused_pass<4, NT, true, SIGN>(X, ldx, A, N, 1 << l, tw, ltw); break;
        case 3: fused_pass<3, NT, true, SIGN>(X, ldx, A, N, 1 << l, tw, ltw); break;
        case 2: fused_pass<2, NT, true, SIGN>(X, ldx, A, N, 1 << l, tw, ltw); break;
        default: fused_pass<1, NT, true, SIGN>(X, ldx, A, N, 1 << l, tw, ltw); break;
      }
      l += K;
    }
  } else {
    for (int l = lN; l > 0;) {  // stages with half-spans 2^(l-1) down to 1
      const int K = l >= 4 ? 4 : l;
      const int h0 = 1 << (l - K);
      switch (K) {
        case 4: fused_pass<4, NT, false, SIGN>(X, ldx, A, N, h0, tw, ltw); break;
        case 3: fused_pass<3, NT, false, SIGN>(X, ldx, A, N, h0, tw, ltw); break;
        case 2: fused_pass<2, NT, false, SIGN>(X, ldx, A, N, h0, tw, ltw); break;
        default: fused_pass<1, NT, false, SIGN>(X, ldx, A, N, h0, tw, ltw); break;
      }
      l -= K;
    }
  }
}

// The middle of a Bluestein convolution as one pass: the last forward DIF
// pass (K stages, half-spans 2^(K-1) .. 1), the product with the chirp
// spectrum (stored bit-reversed, i.e. in the order the DIF leaves the
// values) and the first inverse DIT pass (half-spans 1 .. 2^(K-1)) act on the
// same groups of 2^K consecutive elements, whose twiddles are the constant
// rotations alone; one LDS round trip instead of three.
template <int K, int NT>
__device__ __forceinline__ void mid_pass(cplx<float> *X, int ldx, int A, int L,
                                         const float2 *__restrict__ bspec) {
  constexpr int R = 1 << K;
  const int lg = ilog2(L) - K;
  const int total = A << lg;
  for (int gg = threadIdx.x; gg < total; gg += NT) {
    const int a = gg >> lg, g = gg & ((1 << lg) - 1);
    const int base = g << K;
    cplx<float> *Y = X + a * ldx + pad(base);  // a group never straddles a pad
    cplx<float> v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = Y[q];
#pragma unroll
    for (int s = K - 1; s >= 0; --s) {
      const int d = 1 << s;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        if (q & d) continue;
        const cplx<float> tq = rot16<-1>(cplx<float>{1.0f, 0.0f}, (q & (d - 1)) * (16 >> (s + 1)));
        const cplx<float> x = v[q], y = v[q + d];
        v[q] = cadd(x, y);
        v[q + d] = cmul(csub(x, y), tq);
      }
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const float2 b = bspec[base + q];
      v[q] = cmul(v[q], cplx<float>{b.x, b.y});
    }
#pragma unroll
    for (int s = 0; s < K; ++s) {
      const int d = 1 << s;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        if (q & d) continue;
        const cplx<float> tq = rot16<1>(cplx<float>{1.0f, 0.0f}, (q & (d - 1)) * (16 >> (s + 1)));
        const cplx<float> y = cmul(v[q + d], tq);
        const cplx<float> x = v[q];
        v[q] = cadd(x, y);
        v[q + d] = csub(x, y);
      }
    }
#pragma unroll
    for (int q = 0; q < R; ++q) Y[q] = v[q];
  }
  __syncthreads();
}

// Bluestein convolution of A arrays of length L = 2^lN in place: forward DIF
// passes (K = 4, half-spans down to 2^(l0)), the middle pass (K = l0 <= 4,
// with the chirp spectrum), inverse DIT passes mirroring the forward ones;
// natural order in and out.
template <int NT>
__device__ __forceinline__ void bluestein_conv(cplx<float> *X, int ldx, int A, int L,
                                               const float2 *__restrict__ tw, int twlen,
                                               const float2 *__restrict__ bspec) {
  const int lN = ilog2(L), ltw = ilog2(twlen);
  int l = lN;
  for (; l > 4; l -= 4) fused_pass<4, NT, false, -1>(X, ldx, A, L, 1 << (l - 4), tw, ltw);
  switch (l) {
    case 4: mid_pass<4, NT>(X, ldx, A, L, bspec); break;
    case 3: mid_pass<3, NT>(X, ldx, A, L, bspec); break;
    case 2: mid_pass<2, NT>(X, ldx, A, L, bspec); break;
    default: mid_pass<1, NT>(X, ldx, A, L, bspec); break;
  }
  for (int lo = l; lo < lN; lo += 4) fused_pass<4, NT, true, 1>(X, ldx, A, L, 1 << lo, tw, ltw);
}

// The aliasing fold of one ring's coefficients into its n = 4 r bins, for a
// unit of two shells (h = G(s) + i G(s + 1)), written to the bin slots of X
// (bin k = k1 + r k2 at X[k2 ldx + pad(pos(k1))], pos = bit-reversed k1 on
// lbits bits, or k1 when lbits = 0).  Work item (k, q): the chunks j = q,
// q + Q, ... of m, each holding the direct term m = j n + k and the conjugate
// term m = (j + 1) n - k.  Small rings (n <= NT / 2) get Q = NT / n threads per
// bin, whose partial sums meet in LDS and are added in q order; a bin's direct
// terms are summed before its conjugate terms in both cases.  The loads of
// kBatch chunks (or, when every bin has one chunk, of kBatch bins) are issued
// together at clamped addresses, so a thread keeps many loads in flight.
// MODE: 0 = one shell, 1 = two shells (two 8-byte loads per m), 2 = two
// shells from one 16-byte load (the pair 16-byte aligned).
constexpr int kBatch = 8;

template <int MODE>
__device__ __forceinline__ void load_pair(const cplx<float> *__restrict__ f, cplx<float> &a,
                                          cplx<float> &b) {
  if (MODE == 2) {
    const float4 v = *reinterpret_cast<const float4 *>(f);
    a = cplx<float>{v.x, v.y};
    b = cplx<float>{v.z, v.w};
  } else {
    a = f[0];
    b = MODE == 1 ? f[1] : cplx<float>{0.0f, 0.0f};
  }
}

// The rows m = 0..lmax of one unit (two shells) of the ring buffer through a
// buffer resource whose range ends after the pair of row lmax: a read at
// `oob`, past that end, returns zeros (buffer range checking), so a row that
// does not exist needs neither a branch nor a zero fill.  MODE as load_pair.
template <int MODE>
struct f_rows {
  __amdgpu_buffer_rsrc_t rs;
  uint32_t row, oob;
  __device__ __forceinline__ f_rows(const cplx<float> *Fu, int G, int lmax) {
    row = uint32_t(G) * 8u;
    oob = (uint32_t(lmax) * row + 31u) & ~15u;  // >= the end of row lmax's pair
    rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<cplx<float> *>(Fu), 0, int(oob),
                                           0x00020000);
  }
  __device__ __forceinline__ void load(int m, bool ok, cplx<float> &a, cplx<float> &b) const {
    const uint32_t off = ok ? uint32_t(m) * row : oob;
    if (MODE == 2) {
      const float4 v =
          __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
      a = cplx<float>{v.x, v.y};
      b = cplx<float>{v.z, v.w};
    } else {
      const float2 v = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0));
      a = cplx<float>{v.x, v.y};
      if (MODE == 1) {
        const float2 w =
            __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rs, off + 8u, 0, 0));
        b = cplx<float>{w.x, w.y};
      } else {
        b = cplx<float>{0.0f, 0.0f};
      }
    }
  }
};

template <int NT, int MODE>
__device__ __forceinline__ void fold_bins(const cplx<float> *__restrict__ Fu, int G, int n, int r,
                                          int lmax, int lbits, int ldx, cplx<float> *X) {
  const int t = threadIdx.x;
  const int Q = n <= NT / 2 ? NT / n : 1;
  const int jmax = lmax / n;  // chunks 0..jmax can hold terms
  auto slot = [&](int k) {
    const int k2 = (k >= r) + (k >= 2 * r) + (k >= 3 * r);
    const int k1 = k - k2 * r;
    const int pos = lbits ? int(__brev(unsigned(k1)) >> (32 - lbits)) : k1;
    return k2 * ldx + pad(pos);
  };
  auto pack = [](cplx<float> g0, cplx<float> g1) {
    return cplx<float>{g0.re - g1.im, g0.im + g1.re};
  };
  // terms that do not exist (m > lmax) read zeros: every term is one load and
  // one unconditional addition
  const f_rows<MODE> fr(Fu, G, lmax);
  if (jmax == 0) {
    // n > lmax: a bin has at most one direct and one conjugate term
    for (int k0 = t; k0 < n; k0 += kBatch * NT) {
      cplx<float> fd[kBatch][2], fc[kBatch][2];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int k = min(k0 + u * NT, n - 1);
        fr.load(k, k <= lmax, fd[u][0], fd[u][1]);
        fr.load(n - k, n - k <= lmax, fc[u][0], fc[u][1]);
      }
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int k = k0 + u * NT;
        if (k < n)
          X[slot(k)] = pack(cadd(fd[u][0], cconj(fc[u][0])), cadd(fd[u][1], cconj(fc[u][1])));
      }
    }
    return;
  }
  const int nitems = n * Q;
  cplx<float> hd = {0.0f, 0.0f}, hc = {0.0f, 0.0f};  // the one item of a Q > 1 thread
  for (int it = t; it < nitems; it += NT) {
    const int q = it / n, k = it - q * n;
    cplx<float> d0 = {0.0f, 0.0f}, d1 = d0, c0 = d0, c1 = d0;
    // chunks j0, j0 + Q, ..: full batches of kBatch, then pairs
    auto batch = [&](auto nb, int j0) {
      constexpr int B = decltype(nb)::value;
      cplx<float> fd[B][2], fc[B][2];
#pragma unroll
      for (int u = 0; u < B; ++u) {
        const int j = j0 + u * Q;
        fr.load(j * n + k, j * n + k <= lmax, fd[u][0], fd[u][1]);
        fr.load((j + 1) * n - k, (j + 1) * n - k <= lmax, fc[u][0], fc[u][1]);
      }
#pragma unroll
      for (int u = 0; u < B; ++u) {
        d0 = cadd(d0, fd[u][0]);
        d1 = cadd(d1, fd[u][1]);
        c0 = cadd(c0, cconj(fc[u][0]));
        c1 = cadd(c1, cconj(fc[u][1]));
      }
    };
    int j0 = q;
    for (; j0 + (kBatch - 1) * Q <= jmax; j0 += kBatch * Q)
      batch(std::integral_constant<int, kBatch>{}, j0);
    for (; j0 <= jmax; j0 += 2 * Q) batch(std::integral_constant<int, 2>{}, j0);
    if (Q == 1) {
      X[slot(k)] = pack(cadd(d0, c0), cadd(d1, c1));
    } else {
      hd = pack(d0, d1);
      hc = pack(c0, c1);
    }
  }
  if (Q > 1) {
    // partial sums: direct at X[it], conjugate at X[nitems + it] (nitems <= NT)
    if (t < nitems) {
      X[t] = hd;
      X[nitems + t] = hc;
    }
    __syncthreads();
    cplx<float> sd = {0.0f, 0.0f}, sc = {0.0f, 0.0f};
    if (t < n) {
      for (int q = 0; q < Q; ++q) sd = cadd(sd, X[q * n + t]);
      for (int q = 0; q < Q; ++q) sc = cadd(sc, X[nitems + q * n + t]);
    }
    __syncthreads();
    if (t < n) X[slot(t)] = cadd(sd, sc);
  }
}

// grid: 8 * ceil(nlist / 8) * nunit blocks (XCD-aware as decode_ring_block),
// one unit = one complex transform (two shells); rings taken from ring_list;
// BLU selects the Bluestein form.  All four sub-transforms run together on
// four LDS arrays (one set of twiddles and barriers per pass).
template <int NT, int MAXC, bool BLU>
__global__ __launch_bounds__(NT) void ring_synth2_kernel(
    int lmax, const int32_t *__restrict__ ring_list, int nlist, const int64_t *__restrict__ ring_start,
    const int32_t *__restrict__ ring_n, const int64_t *__restrict__ blu_off,
    const float2 *__restrict__ blu, const float2 *__restrict__ tw, int twlen, int ldx,
    const cplx<float> *__restrict__ F, int S, int G, int nunit, float *__restrict__ maps,
    int64_t npix, const ga_epilogue *__restrict__ epi, const float *__restrict__ ref) {
  extern __shared__ char smem_raw[];
  cplx<float> *X = reinterpret_cast<cplx<float> *>(smem_raw);
  int li, unit;
  if (!decode_ring_block(nlist, nunit, li, unit)) return;
  const int ring = ring_list[li];
  const int t = threadIdx.x;
  const int sbase = 2 * unit;
  const int n = ring_n[ring], r = n >> 2;
  const int ngrp = (S + G - 1) / G;
  const cplx<float> *Fu = F + (int64_t(ring) * ngrp + sbase / G) * (lmax + 1) * G + sbase % G;
  const bool two = sbase + 1 < S;
  const int64_t boff = BLU ? blu_off[r] : 0;
  const float2 *__restrict__ chirp = blu + boff;
  const float2 *__restrict__ bspec = chirp + r;
  const int L = BLU ? blu_len(r) : r;
  const int lbits = ilog2(L);

  // ---- fold into the n bins, h_k = G_k(shell s) + i G_k(shell s + 1), with
  //   G_k = sum_{m = k mod n} F_m + sum_{m = -k mod n, m > 0} conj(F_m)
  // (m <= lmax).  Bin k = k1 + r k2 is kept at X[k2 ldx + pad(pos(k1))], the
  // slot the radix-4 split below reads and rewrites (pos = bit-reversed k1 for
  // the DIT passes of a power-of-two r, k1 for Bluestein).
  const int lb = BLU ? 0 : lbits;
  if (!two)
    fold_bins<NT, 0>(Fu, G, n, r, lmax, lb, ldx, X);
  else if (G % 2 == 0)  // the unit's pair is 16-byte aligned
    fold_bins<NT, 2>(Fu, G, n, r, lmax, lb, ldx, X);
  else
    fold_bins<NT, 1>(Fu, G, n, r, lmax, lb, ldx, X);
  __syncthreads();

  // ---- radix-4 split (thread owns k1 = t + i NT), sub-transform inputs in place
#pragma unroll
  for (int i = 0; i < MAXC; ++i) {
    const int k1 = t + i * NT;
    if (k1 < r) {
      const int pos = BLU ? k1 : (lbits ? int(__brev(unsigned(k1)) >> (32 - lbits)) : 0);
      cplx<float> h[4];
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) h[k2] = X[k2 * ldx + pad(pos)];
      // e^{2 pi i k1 / n}: from the table when n (a power of two) divides its length
      const cplx<float> w1 =
          BLU ? expi_q(k1, n, 0.0f) : twid(tw, k1 << (ilog2(twlen) - ilog2(n)), 1);
      const cplx<float> w2 = cmul(w1, w1);
      const cplx<float> w3 = cmul(w2, w1);
      const cplx<float> s02 = cadd(h[0], h[2]), d02 = csub(h[0], h[2]);
      const cplx<float> s13 = cadd(h[1], h[3]), d13 = cmuli(csub(h[1], h[3]));
      cplx<float> y[4] = {cadd(s02, s13), cmul(cadd(d02, d13), w1), cmul(csub(s02, s13), w2),
                          cmul(csub(d02, d13), w3)};
      if (BLU) {
        const float2 c = chirp[k1];
#pragma unroll
        for (int q = 0; q < 4; ++q) y[q] = cmul(y[q], cplx<float>{c.x, c.y});
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) X[q * ldx + pad(pos)] = y[q];
    }
  }
  if (BLU) {
    for (int e = t; e < (L - r) * 4; e += NT) {
      const int a = e / (L - r), j = r + e - a * (L - r);
      X[a * ldx + pad(j)] = cplx<float>{0.0f, 0.0f};
    }
  }
  __syncthreads();

  // ---- the four length-r inverse DFTs (Bluestein: forward, chirp spectrum, inverse)
  if (BLU) {
    fused_fft<NT, false, -1>(X, ldx, 4, L, tw, twlen);  // natural -> bit-reversed
    for (int e = t; e < L * 4; e += NT) {
      const int a = e >> lbits, j = e & (L - 1);
      const float2 b = bspec[j];  // stored bit-reversed
      X[a * ldx + pad(j)] = cmul(X[a * ldx + pad(j)], cplx<float>{b.x, b.y});
    }
    __syncthreads();
  }
  fused_fft<NT, true, +1>(X, ldx, 4, L, tw, twlen);  // bit-reversed -> natural

  // ---- epilogue: pixels 4 j1 .. 4 j1 + 3 of each shell, one 16-byte store
  const int64_t st = ring_start[ring];
  int k0 = GA_T_NORMAL, kk1 = GA_T_NORMAL;
  float l0 = 1.0f, a0 = 0.0f, l1 = 1.0f, a1 = 0.0f;
  if (epi) {
    k0 = epi[sbase].kind;
    l0 = float(epi[sbase].lamda);
    a0 = float(epi[sbase].param);
    if (two) {
      kk1 = epi[sbase + 1].kind;
      l1 = float(epi[sbase + 1].lamda);
      a1 = float(epi[sbase + 1].param);
    }
  }
  float4 *out0 = reinterpret_cast<float4 *>(maps + int64_t(sbase) * npix + st);
  float4 *out1 = reinterpret_cast<float4 *>(maps + int64_t(sbase + 1) * npix + st);
  for (int j1 = t; j1 < r; j1 += NT) {
    cplx<float> y[4];
    cplx<float> cj = {1.0f, 0.0f};
    if (BLU) {
      const float2 c = chirp[j1];
      cj = cplx<float>{c.x, c.y};
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      y[q] = X[q * ldx + pad(j1)];
      if (BLU) y[q] = cmul(y[q], cj);
    }
    out0[j1] = resid4(ref, int64_t(sbase) * npix + st, j1, float4{apply_epi<float>(y[0].re, k0, l0, a0), apply_epi<float>(y[1].re, k0, l0, a0),
                      apply_epi<float>(y[2].re, k0, l0, a0), apply_epi<float>(y[3].re, k0, l0, a0)});
    if (two)
      out1[j1] = resid4(ref, int64_t(sbase + 1) * npix + st, j1, float4{apply_epi<float>(y[0].im, kk1, l1, a1), apply_epi<float>(y[1].im, kk1, l1, a1),
                        apply_epi<float>(y[2].im, kk1, l1, a1), apply_epi<float>(y[3].im, kk1, l1, a1)});
  }
}

template <int NT, int MAXC, bool BLU>
int launch_synth2(const ga_plan *p, const int32_t *list, int nlist, int N, int S, int G,
                  const float *F, float *maps, const ga_epilogue *epi, const float *ref,
                  hipStream_t stream) {
  if (nlist == 0) return GA_OK;
  const int ldx = padded_len(N);
  // four sub-transform arrays; the fold's partial sums need 2 NT entries
  const size_t lds = size_t(std::max(4 * ldx, 2 * NT)) * sizeof(cplx<float>);
  GA_REQUIRE(lds <= 160 * 1024, GA_ERR_UNSUPPORTED, "ring synthesis needs %zu bytes of LDS", lds);
  const int nunit = (S + 1) / 2;
  const int nblk = 8 * ((nlist + 7) / 8) * nunit;
  auto kfn = ring_synth2_kernel<NT, MAXC, BLU>;
  GA_CHECK_HIP(hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   int(lds)));
  hipLaunchKernelGGL(kfn, dim3(unsigned(nblk)), dim3(NT), lds, stream, p->lmax, list, nlist,
                     p->d_ring_start, p->d_ring_n, p->d_blu_off, p->d_blu32, p->d_tw32,
                     p->tw_len, ldx, reinterpret_cast<const cplx<float> *>(F), S, G, nunit, maps,
                     p->npix, epi, ref);
  GA_LAUNCH_CHECK();
  return GA_OK;
}

// ---------------------------------------------------------------------------
// FP32 synthesis of the wide power-of-two rings (n = 4 r > lmax, r = NT C):
// every bin has at most one direct and one conjugate term, so a thread loads
// the coefficients of its own bins k1 + r k2 (k1 = t + i NT, k2 = 0..3) and
// runs the radix-4 split in registers: no LDS round trip for the fold.  The
// four length-r transforms then run two at a time on two LDS arrays (half
// the LDS of the four-array form, so twice the workgroups per CU): outputs 2
// and 3 of the split wait in registers while arrays 0 and 1 are transformed,
// and the results of 0 and 1 wait in registers while 2 and 3 are.  Same
// arithmetic, term order and passes as ring_synth2_kernel.
// Bin k of a ring with n > lmax holds at most the direct term F_k (k <= lmax)
// and the conjugate term conj(F_{n-k}) (n - k <= lmax).  wide_bin_first loads
// the first present one (conjugated if it is the conjugate term) into g;
// wide_bin_second adds the conjugate term where both are present (bins
// n - lmax <= k <= lmax, few or none), so only one load per bin is in flight.
template <int MODE>
__device__ __forceinline__ void wide_bin_first(const cplx<float> *__restrict__ Fu, int G, int n,
                                               int lmax, int k, bool valid, cplx<float> (&g)[2]) {
  const bool direct = k <= lmax;
  g[0] = g[1] = cplx<float>{0.0f, 0.0f};
  if (valid && (direct || n - k <= lmax))
    load_pair<MODE>(Fu + int64_t(direct ? k : n - k) * G, g[0], g[1]);
  if (!direct) {
    g[0] = cconj(g[0]);
    g[1] = cconj(g[1]);
  }
}
template <int MODE>
__device__ __forceinline__ void wide_bin_second(const cplx<float> *__restrict__ Fu, int G, int n,
                                                int lmax, int k, bool valid, cplx<float> (&g)[2]) {
  if (valid && k <= lmax && n - k <= lmax) {
    cplx<float> c0, c1;
    load_pair<MODE>(Fu + int64_t(n - k) * G, c0, c1);
    g[0] = cadd(g[0], cconj(c0));
    g[1] = cadd(g[1], cconj(c1));
  }
}

template <int NT, int C, int MODE>
__device__ __forceinline__ void fold_split_wide(const cplx<float> *__restrict__ Fu, int G, int r,
                                                int lmax, const float2 *__restrict__ tw,
                                                int ltw, int ldx, cplx<float> *X,
                                                cplx<float> (&park)[C][2]) {
  const int t = threadIdx.x;
  const int n = 4 * r, ln = ilog2(n), lbits = ilog2(r);
  constexpr int B = C >= 2 ? 2 : 1;  // k1 values whose loads are issued together
#pragma unroll
  for (int i0 = 0; i0 < C; i0 += B) {
    cplx<float> g[B][4][2];
#pragma unroll
    for (int u = 0; u < B; ++u)
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2)
        wide_bin_first<MODE>(Fu, G, n, lmax, t + (i0 + u) * NT + k2 * r, true, g[u][k2]);
#pragma unroll
    for (int u = 0; u < B; ++u)
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2)
        wide_bin_second<MODE>(Fu, G, n, lmax, t + (i0 + u) * NT + k2 * r, true, g[u][k2]);
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int k1 = t + (i0 + u) * NT;
      cplx<float> h[4];
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2)
        h[k2] = cplx<float>{g[u][k2][0].re - g[u][k2][1].im, g[u][k2][0].im + g[u][k2][1].re};
      const cplx<float> w1 = twid(tw, k1 << (ltw - ln), 1);
      const cplx<float> w2 = cmul(w1, w1);
      const cplx<float> w3 = cmul(w2, w1);
      const cplx<float> s02 = cadd(h[0], h[2]), d02 = csub(h[0], h[2]);
      const cplx<float> s13 = cadd(h[1], h[3]), d13 = cmuli(csub(h[1], h[3]));
      const int pos = int(__brev(unsigned(k1)) >> (32 - lbits));
      X[pad(pos)] = cadd(s02, s13);
      X[ldx + pad(pos)] = cmul(cadd(d02, d13), w1);
      park[i0 + u][0] = cmul(csub(s02, s13), w2);
      park[i0 + u][1] = cmul(csub(d02, d13), w3);
    }
    __builtin_amdgcn_sched_barrier(0);  // one batch of loads in flight
  }
}

// Paired fold of the wide rings.  Bins k1 + r k2 and (r - k1) + r k2 need
// the same coefficients: n - (k1 + r k2) = (r - k1) + r (3 - k2), so the
// direct terms of one are the conjugate terms of the other.  A thread folds
// both from one set of loads, Fa[j] = F_{a + r j} and Fb[j] = F_{b + r j}
// (zero above lmax; b = r - a):
//   G_{a + r k2} = Fa[k2] + conj(Fb[3 - k2]),  G_{b + r k2} = Fb[k2] + conj(Fa[3 - k2])
// (direct term first, as in wide_bin_first / wide_bin_second).  Thread t owns
// q = t + i NT (i < C / 2, q < (r + 1) / 2) and the outputs k1 = q and
// r - q, in slots 2 i and 2 i + 1; q = 0 owns k1 = 0 (whose conjugate terms
// are F_{r (4 - k2)}: b = r) and, for even r, the self-paired k1 = r / 2.
// Every coefficient of the ring is loaded once, where the unpaired fold
// loads most twice.  The loads of B values of q are issued together: j = 0,
// 1 (every term when lmax < 2 r, all but F_{2r} at lmax = 2 r), then, only
// where lmax >= 2 r, j = 2, 3.  sink(slot, k1, valid, h) takes the folded
// bins h[k2] of output k1, the two shells of the unit packed as one complex
// transform.
__device__ __forceinline__ int wide_pair_k1(int q, int r, int second) {
  return second ? (q ? r - q : r / 2) : q;
}
__device__ __forceinline__ bool wide_pair_valid(int q, int r, int second) {
  return q < (r + 1) / 2 && (!second || q || !(r & 1));
}

template <int NT, int C, int B, int MODE, typename Sink>
__device__ __forceinline__ void fold_wide_pairs(const cplx<float> *__restrict__ Fu, int G, int r,
                                                int lmax, Sink &&sink) {
  static_assert(C % 2 == 0 && (C / 2) % B == 0, "paired fold: C / 2 q values in batches of B");
  const int t = threadIdx.x;
  constexpr int P = C / 2;
  const bool far = lmax >= 2 * r;  // terms at j >= 2 (block-uniform)
  const f_rows<MODE> fr(Fu, G, lmax);
#pragma unroll
  for (int i0 = 0; i0 < P; i0 += B) {
    cplx<float> f[B][2][4][2];  // [q][side a / b][j][shell of the pair]
    // every term is one load; terms that do not exist read zeros
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int q = t + (i0 + u) * NT;
      const bool okq = q < (r + 1) / 2;
      const int ab[2] = {q, q ? r - q : r};
#pragma unroll
      for (int sd = 0; sd < 2; ++sd)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int mm = ab[sd] + r * j;
          fr.load(mm, (j < 2 || far) && okq && mm <= lmax, f[u][sd][j][0], f[u][sd][j][1]);
        }
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int q = t + (i0 + u) * NT;
#pragma unroll
      for (int second = 0; second < 2; ++second) {
        const bool valid = wide_pair_valid(q, r, second);
        const int k1 = wide_pair_k1(q, r, second);
        cplx<float> g[4][2];
        if (second && q == 0) {
          // k1 = r / 2 (even r): its own coefficients F_{r/2 + r j}, paired with itself
          cplx<float> c[4][2];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int mm = r / 2 + r * j;
            fr.load(mm, valid && mm <= lmax, c[j][0], c[j][1]);
          }
#pragma unroll
          for (int k2 = 0; k2 < 4; ++k2)
#pragma unroll
            for (int h = 0; h < 2; ++h) g[k2][h] = cadd(c[k2][h], cconj(c[3 - k2][h]));
        } else {
#pragma unroll
          for (int k2 = 0; k2 < 4; ++k2)
#pragma unroll
            for (int h = 0; h < 2; ++h)
              g[k2][h] = cadd(f[u][second][k2][h], cconj(f[u][1 - second][3 - k2][h]));
        }
        cplx<float> hh[4];
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2)
          hh[k2] = cplx<float>{g[k2][0].re - g[k2][1].im, g[k2][0].im + g[k2][1].re};
        sink(2 * (i0 + u) + second, k1, valid, hh);
      }
    }
    __builtin_amdgcn_sched_barrier(0);  // one batch of loads in flight
  }
}

// the paired fold of a wide power-of-two ring and its radix-4 split:
// sub-transform inputs 0 and 1 to LDS (bit-reversed), 2 and 3 to park
template <int NT, int C, int MODE>
__device__ __forceinline__ void fold_split_wide_paired(const cplx<float> *__restrict__ Fu, int G,
                                                       int r, int lmax,
                                                       const float2 *__restrict__ tw, int ltw,
                                                       int ldx, cplx<float> *X,
                                                       cplx<float> (&park)[C][2]) {
  const int n = 4 * r, ln = ilog2(n), lbits = ilog2(r);
  fold_wide_pairs<NT, C, (C / 2 >= 2 ? 2 : 1), MODE>(
      Fu, G, r, lmax, [&](int slot, int k1, bool, const cplx<float> (&hh)[4]) {
        const cplx<float> w1 = twid(tw, k1 << (ltw - ln), 1);
        const cplx<float> w2 = cmul(w1, w1);
        const cplx<float> w3 = cmul(w2, w1);
        const cplx<float> s02 = cadd(hh[0], hh[2]), d02 = csub(hh[0], hh[2]);
        const cplx<float> s13 = cadd(hh[1], hh[3]), d13 = cmuli(csub(hh[1], hh[3]));
        const int pos = int(__brev(unsigned(k1)) >> (32 - lbits));
        X[pad(pos)] = cadd(s02, s13);
        X[ldx + pad(pos)] = cmul(cadd(d02, d13), w1);
        park[slot][0] = cmul(csub(s02, s13), w2);
        park[slot][1] = cmul(csub(d02, d13), w3);
      });
}

template <int NT, int C>
__global__ __launch_bounds__(NT, 4) void ring_synth_wide_kernel(
    int lmax, const int32_t *__restrict__ ring_list, int nlist,
    const int64_t *__restrict__ ring_start, int r, const float2 *__restrict__ tw, int twlen,
    int ldx, const cplx<float> *__restrict__ F, int S, int G, int nunit,
    float *__restrict__ maps, int64_t npix, const ga_epilogue *__restrict__ epi,
    const float *__restrict__ ref) {
  extern __shared__ char smem_raw[];
  cplx<float> *X = reinterpret_cast<cplx<float> *>(smem_raw);
  int li, unit;
  if (!decode_ring_block(nlist, nunit, li, unit)) return;
  const int ring = ring_list[li];
  const int t = threadIdx.x;
  const int sbase = 2 * unit;
  const int ngrp = (S + G - 1) / G;
  const cplx<float> *Fu = F + (int64_t(ring) * ngrp + sbase / G) * (lmax + 1) * G + sbase % G;
  const bool two = sbase + 1 < S;
  const int ltw = ilog2(twlen), lbits = ilog2(r);

  // ---- fold + radix-4 split; sub-transform inputs 0, 1 to LDS (bit-reversed)
  // (one k1 per thread, C = 1: the unpaired fold)
  cplx<float> park[C][2];
  if constexpr (C % 2 == 0) {
    if (!two)
      fold_split_wide_paired<NT, C, 0>(Fu, G, r, lmax, tw, ltw, ldx, X, park);
    else if (G % 2 == 0)  // the unit's pair is 16-byte aligned
      fold_split_wide_paired<NT, C, 2>(Fu, G, r, lmax, tw, ltw, ldx, X, park);
    else
      fold_split_wide_paired<NT, C, 1>(Fu, G, r, lmax, tw, ltw, ldx, X, park);
  } else {
    if (!two)
      fold_split_wide<NT, C, 0>(Fu, G, r, lmax, tw, ltw, ldx, X, park);
    else if (G % 2 == 0)
      fold_split_wide<NT, C, 2>(Fu, G, r, lmax, tw, ltw, ldx, X, park);
    else
      fold_split_wide<NT, C, 1>(Fu, G, r, lmax, tw, ltw, ldx, X, park);
  }
  __syncthreads();
  fused_fft<NT, true, +1>(X, ldx, 2, r, tw, twlen);
  // ---- results of 0, 1 to registers; inputs of 2, 3 to LDS
  cplx<float> res[C][2];
#pragma unroll
  for (int i = 0; i < C; ++i) {
    const int j1 = t + i * NT;
    res[i][0] = X[pad(j1)];
    res[i][1] = X[ldx + pad(j1)];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < C; ++i) {
    const int k1 = C % 2 == 0 ? wide_pair_k1(t + (i / 2) * NT, r, i & 1) : t + i * NT;
    const int pos = int(__brev(unsigned(k1)) >> (32 - lbits));
    X[pad(pos)] = park[i][0];
    X[ldx + pad(pos)] = park[i][1];
  }
  __syncthreads();
  fused_fft<NT, true, +1>(X, ldx, 2, r, tw, twlen);

  // ---- epilogue: pixels 4 j1 .. 4 j1 + 3 of each shell, one 16-byte store
  const int64_t st = ring_start[ring];
  int k0 = GA_T_NORMAL, kk1 = GA_T_NORMAL;
  float l0 = 1.0f, a0 = 0.0f, l1 = 1.0f, a1 = 0.0f;
  if (epi) {
    k0 = epi[sbase].kind;
    l0 = float(epi[sbase].lamda);
    a0 = float(epi[sbase].param);
    if (two) {
      kk1 = epi[sbase + 1].kind;
      l1 = float(epi[sbase + 1].lamda);
      a1 = float(epi[sbase + 1].param);
    }
  }
  float4 *out0 = reinterpret_cast<float4 *>(maps + int64_t(sbase) * npix + st);
  float4 *out1 = reinterpret_cast<float4 *>(maps + int64_t(sbase + 1) * npix + st);
#pragma unroll
  for (int i = 0; i < C; ++i) {
    const int j1 = t + i * NT;
    const cplx<float> y[4] = {res[i][0], res[i][1], X[pad(j1)], X[ldx + pad(j1)]};
    out0[j1] = resid4(ref, int64_t(sbase) * npix + st, j1, float4{apply_epi<float>(y[0].re, k0, l0, a0), apply_epi<float>(y[1].re, k0, l0, a0),
                      apply_epi<float>(y[2].re, k0, l0, a0), apply_epi<float>(y[3].re, k0, l0, a0)});
    if (two)
      out1[j1] = resid4(ref, int64_t(sbase + 1) * npix + st, j1, float4{apply_epi<float>(y[0].im, kk1, l1, a1), apply_epi<float>(y[1].im, kk1, l1, a1),
                        apply_epi<float>(y[2].im, kk1, l1, a1), apply_epi<float>(y[3].im, kk1, l1, a1)});
  }
}

template <int NT, int C>
int launch_synth_wide(const ga_plan *p, int S, int G, const float *F, float *maps,
                      const ga_epilogue *epi, const float *ref, hipStream_t stream) {
  const int r = p->r_wide;
  GA_REQUIRE(r == NT * C, GA_ERR_UNSUPPORTED, "wide ring kernel for r = %d launched with r = %d",
             NT * C, r);
  const int ldx = padded_len(r);
  const size_t lds = size_t(2) * ldx * sizeof(cplx<float>);
  const int nunit = (S + 1) / 2;
  const int nblk = 8 * ((p->n_wide + 7) / 8) * nunit;
  auto kfn = ring_synth_wide_kernel<NT, C>;
  GA_CHECK_HIP(hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   int(lds)));
  hipLaunchKernelGGL(kfn, dim3(unsigned(nblk)), dim3(NT), lds, stream, p->lmax, p->d_rings_wide,
                     p->n_wide, p->d_ring_start, r, p->d_tw32, p->tw_len, ldx,
                     reinterpret_cast<const cplx<float> *>(F), S, G, nunit, maps, p->npix, epi,
                     ref);
  GA_LAUNCH_CHECK();
  return GA_OK;
}

// ---------------------------------------------------------------------------
// FP32 synthesis of the wide Bluestein rings (non-power-of-two r, n = 4 r >
// lmax, L = blu_len(r) <= Lmax of the class): the fold and the radix-4 split
// run in registers as for the wide power-of-two rings (bins k1 + r k2 of the
// thread's k1 = t + c NT, c < C), then the four Bluestein convolutions run A
// at a time on A LDS arrays of Lmax (A = 2: half the LDS of the four-array
// form): the remaining split outputs and the finished results
// wait in registers.  Launched with A = 2.  Same arithmetic, term order and
// passes as ring_synth2_kernel with BLU.

// radix-4 split of the bins h[k2] = bin k1 + r k2 of a Bluestein ring and the
// input chirp: y[q] = (sum_k2 i^(q k2) h[k2]) e^{2 pi i q k1 / n} chirp[k1]
__device__ __forceinline__ void split_chirp(const cplx<float> (&h)[4], int k1, int n,
                                            const float2 *__restrict__ chirp,
                                            cplx<float> (&y)[4]) {
  const cplx<float> w1 = expi_q(k1, n, 0.0f);
  const cplx<float> w2 = cmul(w1, w1);
  const cplx<float> w3 = cmul(w2, w1);
  const cplx<float> s02 = cadd(h[0], h[2]), d02 = csub(h[0], h[2]);
  const cplx<float> s13 = cadd(h[1], h[3]), d13 = cmuli(csub(h[1], h[3]));
  const float2 cc = chirp[k1];
  const cplx<float> ch = {cc.x, cc.y};
  y[0] = cmul(cadd(s02, s13), ch);
  y[1] = cmul(cmul(cadd(d02, d13), w1), ch);
  y[2] = cmul(cmul(csub(s02, s13), w2), ch);
  y[3] = cmul(cmul(csub(d02, d13), w3), ch);
}

template <int NT, int C, int MODE>
__device__ __forceinline__ void fold_split_wide_blu(const cplx<float> *__restrict__ Fu, int G,
                                                    int r, int lmax,
                                                    const float2 *__restrict__ chirp,
                                                    cplx<float> (&y)[C][4]) {
  const int t = threadIdx.x;
  const int n = 4 * r;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int k1 = t + c * NT;
    cplx<float> g[4][2];
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) wide_bin_first<MODE>(Fu, G, n, lmax, k1 + k2 * r, k1 < r, g[k2]);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) wide_bin_second<MODE>(Fu, G, n, lmax, k1 + k2 * r, k1 < r, g[k2]);
    cplx<float> h[4];
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2)
      h[k2] = cplx<float>{g[k2][0].re - g[k2][1].im, g[k2][0].im + g[k2][1].re};
    split_chirp(h, k1 < r ? k1 : 0, n, chirp, y[c]);
    // loads of at most two k1 values in flight (registers)
    if (c % 2 == 1) __builtin_amdgcn_sched_barrier(0);
  }
}

// the same from bins already folded into LDS (bin k1 + r k2 at X[k2 ldf + pad(k1)])
template <int NT, int C>
__device__ __forceinline__ void split_from_lds(const cplx<float> *X, int ldf, int r,
                                               const float2 *__restrict__ chirp,
                                               cplx<float> (&y)[C][4]) {
  const int n = 4 * r;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int k1 = min(int(threadIdx.x) + c * NT, r - 1);
    cplx<float> h[4];
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) h[k2] = X[k2 * ldf + pad(k1)];
    split_chirp(h, k1, n, chirp, y[c]);
  }
}

template <int NT, int C, int A, bool LDSFOLD>
__global__ __launch_bounds__(NT, (NT == 256 && C == 4) ? 3 : 4) void ring_synth_wide_blu_kernel(
    int lmax, const int32_t *__restrict__ ring_list, int nlist,
    const int64_t *__restrict__ ring_start, const int32_t *__restrict__ ring_n,
    const int64_t *__restrict__ blu_off, const float2 *__restrict__ blu,
    const float2 *__restrict__ tw, int twlen, int ldx, const cplx<float> *__restrict__ F, int S,
    int G, int nunit, float *__restrict__ maps, int64_t npix,
    const ga_epilogue *__restrict__ epi, const float *__restrict__ ref) {
  extern __shared__ char smem_raw[];
  cplx<float> *X = reinterpret_cast<cplx<float> *>(smem_raw);
  int li, unit;
  if (!decode_ring_block(nlist, nunit, li, unit)) return;
  const int ring = ring_list[li];
  const int t = threadIdx.x;
  const int sbase = 2 * unit;
  const int r = ring_n[ring] >> 2;
  const int ngrp = (S + G - 1) / G;
  const cplx<float> *Fu = F + (int64_t(ring) * ngrp + sbase / G) * (lmax + 1) * G + sbase % G;
  const bool two = sbase + 1 < S;
  const float2 *__restrict__ chirp = blu + blu_off[r];
  const float2 *__restrict__ bspec = chirp + r;
  const int L = blu_len(r), lbits = ilog2(L);

  // ---- fold + radix-4 split + input chirp, in registers (LDSFOLD: the fold
  // of a ring with n <= lmax goes through four packed LDS arrays first)
  cplx<float> y[C][4];
  if (LDSFOLD) {
    const int n = 4 * r, ldf = padded_len(r);
    if (!two)
      fold_bins<NT, 0>(Fu, G, n, r, lmax, 0, ldf, X);
    else if (G % 2 == 0)
      fold_bins<NT, 2>(Fu, G, n, r, lmax, 0, ldf, X);
    else
      fold_bins<NT, 1>(Fu, G, n, r, lmax, 0, ldf, X);
    __syncthreads();
    split_from_lds<NT, C>(X, ldf, r, chirp, y);
    __syncthreads();
  } else if constexpr (C % 2 == 0) {
    // paired fold (slot c holds output k1 = blu_slot_k1(c))
    const int n = 4 * r;
    auto sink = [&](int slot, int k1, bool valid, const cplx<float> (&hh)[4]) {
      split_chirp(hh, valid ? k1 : 0, n, chirp, y[slot]);
    };
    if (!two)
      fold_wide_pairs<NT, C, 1, 0>(Fu, G, r, lmax, sink);
    else if (G % 2 == 0)  // the unit's pair is 16-byte aligned
      fold_wide_pairs<NT, C, 1, 2>(Fu, G, r, lmax, sink);
    else
      fold_wide_pairs<NT, C, 1, 1>(Fu, G, r, lmax, sink);
  } else {
    if (!two)
      fold_split_wide_blu<NT, C, 0>(Fu, G, r, lmax, chirp, y);
    else if (G % 2 == 0)  // the unit's pair is 16-byte aligned
      fold_split_wide_blu<NT, C, 2>(Fu, G, r, lmax, chirp, y);
    else
      fold_split_wide_blu<NT, C, 1>(Fu, G, r, lmax, chirp, y);
  }
  // input position of slot c: k1 = t + c NT, or the paired fold's output
  auto in_k1 = [&](int c, bool &valid) {
    if (!LDSFOLD && C % 2 == 0) {
      const int q = t + (c / 2) * NT;
      valid = wide_pair_valid(q, r, c & 1);
      return wide_pair_k1(q, r, c & 1);
    }
    valid = t + c * NT < r;
    return t + c * NT;
  };

  // ---- the four convolutions, A at a time; results back into y
#pragma unroll
  for (int q0 = 0; q0 < 4; q0 += A) {
#pragma unroll
    for (int a = 0; a < A; ++a)
#pragma unroll
      for (int c = 0; c < C; ++c) {
        bool valid;
        const int k1 = in_k1(c, valid);
        if (valid) X[a * ldx + pad(k1)] = y[c][q0 + a];
      }
    for (int e = t; e < (L - r) * A; e += NT) {
      const int a = e / (L - r), j = r + e - a * (L - r);
      X[a * ldx + pad(j)] = cplx<float>{0.0f, 0.0f};
    }
    __syncthreads();
    bluestein_conv<NT>(X, ldx, A, L, tw, twlen, bspec);
#pragma unroll
    for (int a = 0; a < A; ++a)
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int j1 = t + c * NT;
        if (j1 < r) y[c][q0 + a] = X[a * ldx + pad(j1)];
      }
    __syncthreads();
  }

  // ---- epilogue: output chirp, pixels 4 j1 .. 4 j1 + 3 of each shell
  const int64_t st = ring_start[ring];
  int k0 = GA_T_NORMAL, kk1 = GA_T_NORMAL;
  float l0 = 1.0f, a0 = 0.0f, l1 = 1.0f, a1 = 0.0f;
  if (epi) {
    k0 = epi[sbase].kind;
    l0 = float(epi[sbase].lamda);
    a0 = float(epi[sbase].param);
    if (two) {
      kk1 = epi[sbase + 1].kind;
      l1 = float(epi[sbase + 1].lamda);
      a1 = float(epi[sbase + 1].param);
    }
  }
  float4 *out0 = reinterpret_cast<float4 *>(maps + int64_t(sbase) * npix + st);
  float4 *out1 = reinterpret_cast<float4 *>(maps + int64_t(sbase + 1) * npix + st);
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int j1 = t + c * NT;
    if (j1 < r) {
      const float2 cc = chirp[j1];
      const cplx<float> cj = {cc.x, cc.y};
      cplx<float> v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = cmul(y[c][q], cj);
      out0[j1] = resid4(ref, int64_t(sbase) * npix + st, j1, float4{apply_epi<float>(v[0].re, k0, l0, a0), apply_epi<float>(v[1].re, k0, l0, a0),
                        apply_epi<float>(v[2].re, k0, l0, a0), apply_epi<float>(v[3].re, k0, l0, a0)});
      if (two)
        out1[j1] = resid4(ref, int64_t(sbase + 1) * npix + st, j1, float4{apply_epi<float>(v[0].im, kk1, l1, a1), apply_epi<float>(v[1].im, kk1, l1, a1),
                          apply_epi<float>(v[2].im, kk1, l1, a1), apply_epi<float>(v[3].im, kk1, l1, a1)});
    }
  }
}

// one Bluestein class (list, count, class maxima of r and L) by the
// register-split kernel; LDSFOLD for rings with n <= lmax
template <int NT, int C, int A, bool LDSFOLD>
int launch_synth_wide_blu(const ga_plan *p, const int32_t *list, int nlist, int max_r, int max_L,
                          int S, int G, const float *F, float *maps, const ga_epilogue *epi,
                          const float *ref, hipStream_t stream) {
  GA_REQUIRE(max_r <= NT * C, GA_ERR_UNSUPPORTED, "Bluestein kernel: r %d above %d", max_r,
             NT * C);
  const int ldx = padded_len(max_L);
  size_t ent = size_t(A) * ldx;
  // the LDS fold: four packed arrays of r, and 2 NT partial sums of small rings
  if (LDSFOLD) ent = std::max({ent, size_t(4) * padded_len(max_r), size_t(2) * NT});
  const size_t lds = ent * sizeof(cplx<float>);
  GA_REQUIRE(lds <= 160 * 1024, GA_ERR_UNSUPPORTED, "ring synthesis needs %zu bytes of LDS", lds);
  const int nunit = (S + 1) / 2;
  const int nblk = 8 * ((nlist + 7) / 8) * nunit;
  auto kfn = ring_synth_wide_blu_kernel<NT, C, A, LDSFOLD>;
  GA_CHECK_HIP(hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   int(lds)));
  hipLaunchKernelGGL(kfn, dim3(unsigned(nblk)), dim3(NT), lds, stream, p->lmax, list, nlist,
                     p->d_ring_start, p->d_ring_n, p->d_blu_off, p->d_blu32, p->d_tw32,
                     p->tw_len, ldx, reinterpret_cast<const cplx<float> *>(F), S, G, nunit, maps,
                     p->npix, epi, ref);
  GA_LAUNCH_CHECK();
  return GA_OK;
}

template <bool LDSFOLD>
int launch_blu_class(const ga_plan *p, const int32_t *list, int nlist, int max_r, int max_L,
                     int S, int G, const float *F, float *maps, const ga_epilogue *epi,
                     const float *ref, hipStream_t stream) {
  // two convolutions at a time; at most four k1 values per thread
  const int cr = (max_r + 255) / 256;
  if (cr <= 1)
    return launch_synth_wide_blu<256, 1, 2, LDSFOLD>(p, list, nlist, max_r, max_L, S, G, F, maps,
                                                     epi, ref, stream);
  if (cr <= 2)
    return launch_synth_wide_blu<256, 2, 2, LDSFOLD>(p, list, nlist, max_r, max_L, S, G, F, maps,
                                                     epi, ref, stream);
  if (cr <= 4)
    return launch_synth_wide_blu<256, 4, 2, LDSFOLD>(p, list, nlist, max_r, max_L, S, G, F, maps,
                                                     epi, ref, stream);
  return launch_synth_wide_blu<512, 4, 2, LDSFOLD>(p, list, nlist, max_r, max_L, S, G, F, maps,
                                                   epi, ref, stream);
}

// ---------------------------------------------------------------------------
// FP32 analysis (pixels -> ring coefficients), register form.  The mirror of
// the synthesis: x_{4 j1 + j2} = p_s0 + i p_s1 (thread owns j1 = t + c NT,
// one 16-byte load per shell), four length-r forward DFTs
// D_{j2}[k1] = sum_{j1} x_{4 j1 + j2} e^{-2 pi i j1 k1 / r} two at a time on
// two LDS arrays (power of two: forward DIF passes, results read at the
// bit-reversed slots; Bluestein: D = conj(c IDFT(conj(x) c)) with the
// synthesis convolution), then in registers
//   X_{k1 + r k2} = sum_{j2} (-i)^{j2 k2} e^{-2 pi i j2 k1 / n} D_{j2}[k1].
// The two real transforms are split with the partner bin n - k, exchanged
// through LDS in two halves (bins k2 = 2, 3 first, for the outputs of
// k2 = 0, 1; then the other way), and every m = k (mod n), m <= lmax, gets
// weight e^{-i m phi0} (d0, d1) stored as one 16-byte pair.
template <int NT, int C, bool BLU>
__global__ __launch_bounds__(NT, (NT == 256 && C >= 4) ? 3 : 4) void ring_anal2_kernel(
    int lmax, const int32_t *__restrict__ ring_list, int nlist,
    const int64_t *__restrict__ ring_start, const int32_t *__restrict__ ring_n,
    const int32_t *__restrict__ ring_pnum, const int32_t *__restrict__ ring_pden,
    const int64_t *__restrict__ blu_off, const float2 *__restrict__ blu,
    const float2 *__restrict__ tw, int twlen, int ldx, const float *__restrict__ maps, int S,
    int nunit, int64_t npix, float weight, cplx<float> *__restrict__ F) {
  extern __shared__ char smem_raw[];
  cplx<float> *X = reinterpret_cast<cplx<float> *>(smem_raw);
  int li, unit;
  if (!decode_ring_block(nlist, nunit, li, unit)) return;
  const int ring = ring_list[li];
  const int t = threadIdx.x;
  const int s0 = 2 * unit;
  const bool two = s0 + 1 < S;
  const int n = ring_n[ring], r = n >> 2;
  const int pnum = ring_pnum[ring], pden = ring_pden[ring];
  const int64_t st = ring_start[ring];
  const float2 *__restrict__ chirp = BLU ? blu + blu_off[r] : blu;
  const float2 *__restrict__ bspec = chirp + r;
  const int L = BLU ? blu_len(r) : r;
  const int lbits = ilog2(r);

  // ---- pixels
  cplx<float> x[C][4];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int j1 = t + c * NT;
    float4 a = float4{0.0f, 0.0f, 0.0f, 0.0f}, b = a;
    if (j1 < r) {
      a = reinterpret_cast<const float4 *>(maps + int64_t(s0) * npix + st)[j1];
      if (two) b = reinterpret_cast<const float4 *>(maps + int64_t(s0 + 1) * npix + st)[j1];
    }
    x[c][0] = cplx<float>{a.x, b.x};
    x[c][1] = cplx<float>{a.y, b.y};
    x[c][2] = cplx<float>{a.z, b.z};
    x[c][3] = cplx<float>{a.w, b.w};
  }

  // ---- the four forward DFTs, two at a time; results back into x
#pragma unroll
  for (int q0 = 0; q0 < 4; q0 += 2) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int j1 = t + c * NT;
        if (j1 < r) {
          // q0 is 0 or 2 (this loop is not unrolled): static indices keep x in registers
          cplx<float> v = q0 == 0 ? x[c][a] : x[c][2 + a];
          if (BLU) {
            const float2 cc = chirp[j1];
            v = cmul(cconj(v), cplx<float>{cc.x, cc.y});
          }
          X[a * ldx + pad(j1)] = v;
        }
      }
    if (BLU) {
      for (int e = t; e < (L - r) * 2; e += NT) {
        const int a = e / (L - r), j = r + e - a * (L - r);
        X[a * ldx + pad(j)] = cplx<float>{0.0f, 0.0f};
      }
    }
    __syncthreads();
    if (BLU)
      bluestein_conv<NT>(X, ldx, 2, L, tw, twlen, bspec);
    else
      fused_fft<NT, false, -1>(X, ldx, 2, r, tw, twlen);  // natural -> bit-reversed
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int k1 = t + c * NT;
        if (k1 < r) {
          cplx<float> v;
          if (BLU) {
            const float2 cc = chirp[k1];
            v = cconj(cmul(X[a * ldx + pad(k1)], cplx<float>{cc.x, cc.y}));
          } else {
            const int pos = lbits ? int(__brev(unsigned(k1)) >> (32 - lbits)) : 0;
            v = X[a * ldx + pad(pos)];
          }
          if (q0 == 0)
            x[c][a] = v;
          else
            x[c][2 + a] = v;
        }
      }
    __syncthreads();
  }

  // ---- radix-4 combination: bins k1 + r k2 into x[c][k2]
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int k1 = min(t + c * NT, r - 1);
    const cplx<float> w1 = cconj(expi_q(k1, n, 0.0f));
    const cplx<float> w2 = cmul(w1, w1), w3 = cmul(w2, w1);
    const cplx<float> y0 = x[c][0], y1 = cmul(x[c][1], w1), y2 = cmul(x[c][2], w2),
                      y3 = cmul(x[c][3], w3);
    const cplx<float> s02 = cadd(y0, y2), d02 = csub(y0, y2);
    const cplx<float> s13 = cadd(y1, y3), d13 = cmuli(csub(y3, y1));  // -i (y1 - y3)
    x[c][0] = cadd(s02, s13);
    x[c][1] = cadd(d02, d13);
    x[c][2] = csub(s02, s13);
    x[c][3] = csub(d02, d13);
  }

  // ---- split the two real transforms and unfold, in two halves: the
  // partner of bin k1 + r k2 is n - k = (r - k1) + r (3 - k2) for k1 > 0,
  // r ((4 - k2) mod 4) for k1 = 0
  cplx<float> *Fr = F + int64_t(ring) * (lmax + 1) * S + s0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    // LDS holds the other half: bins k2 = 2 - 2h + a at X[a ldx + pad(k1)]
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int k1 = t + c * NT;
      if (k1 < r) {  // h is 0 or 1 (this loop is not unrolled): static indices
        X[pad(k1)] = h == 0 ? x[c][2] : x[c][0];
        X[ldx + pad(k1)] = h == 0 ? x[c][3] : x[c][1];
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int k1 = t + c * NT;
      if (k1 >= r) continue;
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int k2 = 2 * h + b;
        const int k = k1 + r * k2;
        if (k > lmax) continue;
        // partner bin (p1, p2): in the other half, or the bin itself (k = 0, n/2)
        const int p1 = k1 ? r - k1 : 0, p2 = k1 ? 3 - k2 : (4 - k2) & 3;
        const cplx<float> xv = h == 0 ? x[c][b] : x[c][2 + b];  // x[c][k2]
        const cplx<float> pv = p2 == k2 ? xv : X[(p2 - (2 - 2 * h)) * ldx + pad(p1)];
        const cplx<float> av = xv, bv = cconj(pv);
        const cplx<float> d0 = {0.5f * (av.re + bv.re), 0.5f * (av.im + bv.im)};
        const cplx<float> d1 = {0.5f * (av.im - bv.im), -0.5f * (av.re - bv.re)};  // (a - b) / (2i)
        // ring phase e^{-i m phi0}, m phi0 = 2 pi (m pnum mod pden) / pden, stepped by n
        int q = pnum ? int(int64_t(k) * pnum % pden) : 0;
        const int dq = pnum ? int(int64_t(n) * pnum % pden) : 0;
        for (int mm = k; mm <= lmax; mm += n) {
          const cplx<float> ph = pnum ? cconj(expi_q(q, pden, 0.0f)) : cplx<float>{1.0f, 0.0f};
          q += dq;
          if (q >= pden) q -= pden;
          const cplx<float> v0 = cmul(d0, ph), v1 = cmul(d1, ph);
          cplx<float> *f = Fr + int64_t(mm) * S;
          if (two && !(S & 1)) {  // the pair is 16-byte aligned
            *reinterpret_cast<float4 *>(f) =
                float4{weight * v0.re, weight * v0.im, weight * v1.re, weight * v1.im};
          } else {
            f[0] = cplx<float>{weight * v0.re, weight * v0.im};
            if (two) f[1] = cplx<float>{weight * v1.re, weight * v1.im};
          }
        }
      }
    }
    __syncthreads();
  }
}

template <int NT, int C, bool BLU>
int launch_anal2(const ga_plan *p, const int32_t *list, int nlist, int maxN, int S,
                 const float *maps, float *F, hipStream_t stream) {
  if (nlist == 0) return GA_OK;
  const int ldx = padded_len(maxN);
  const size_t lds = size_t(2) * ldx * sizeof(cplx<float>);
  GA_REQUIRE(lds <= 160 * 1024, GA_ERR_UNSUPPORTED, "ring analysis needs %zu bytes of LDS", lds);
  const int nunit = (S + 1) / 2;
  const int nblk = 8 * ((nlist + 7) / 8) * nunit;
  auto kfn = ring_anal2_kernel<NT, C, BLU>;
  GA_CHECK_HIP(hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   int(lds)));
  hipLaunchKernelGGL(kfn, dim3(unsigned(nblk)), dim3(NT), lds, stream, p->lmax, list, nlist,
                     p->d_ring_start, p->d_ring_n, p->d_ring_pnum, p->d_ring_pden, p->d_blu_off,
                     p->d_blu32, p->d_tw32, p->tw_len, ldx, maps, S, nunit, p->npix,
                     float(4.0 * M_PI / double(p->npix)), reinterpret_cast<cplx<float> *>(F));
  GA_LAUNCH_CHECK();
  return GA_OK;
}

// both ring classes by the register form; GA_ERR_UNSUPPORTED (before any
// launch) outside its sizes or alignment
int launch_ring_anal2_f32(const ga_plan *p, int S, const float *maps, float *F,
                          hipStream_t stream) {
  if (reinterpret_cast<uintptr_t>(maps) & 15) return GA_ERR_UNSUPPORTED;
  if (S > 1 && (reinterpret_cast<uintptr_t>(F) & 15)) return GA_ERR_UNSUPPORTED;
  const int rp = p->max_r_pow2, rb = p->max_r_blu, Lb = p->max_L_blu;
  if (p->n_pow2 > 0 && rp > 4096) return GA_ERR_UNSUPPORTED;
  if (p->n_blu > 0 && (rb > 2048 || Lb > 4096)) return GA_ERR_UNSUPPORTED;
  int rc = GA_OK;
  if (p->n_pow2 > 0) {
    const int32_t *l = p->d_rings_pow2;
    const int n = p->n_pow2;
    if (rp <= 256) rc = launch_anal2<256, 1, false>(p, l, n, rp, S, maps, F, stream);
    else if (rp <= 512) rc = launch_anal2<256, 2, false>(p, l, n, rp, S, maps, F, stream);
    else if (rp <= 1024) rc = launch_anal2<256, 4, false>(p, l, n, rp, S, maps, F, stream);
    else if (rp <= 2048) rc = launch_anal2<256, 8, false>(p, l, n, rp, S, maps, F, stream);
    else rc = launch_anal2<512, 8, false>(p, l, n, rp, S, maps, F, stream);
    if (rc) return rc;
  }
  if (p->n_blu > 0) {
    const int32_t *l = p->d_rings_blu;
    const int n = p->n_blu;
    if (rb <= 256) rc = launch_anal2<256, 1, true>(p, l, n, Lb, S, maps, F, stream);
    else if (rb <= 512) rc = launch_anal2<256, 2, true>(p, l, n, Lb, S, maps, F, stream);
    else if (rb <= 1024) rc = launch_anal2<256, 4, true>(p, l, n, Lb, S, maps, F, stream);
    else rc = launch_anal2<512, 4, true>(p, l, n, Lb, S, maps, F, stream);
  }
  return rc;
}

// both ring classes with the fused-pass kernels (the wide power-of-two rings
// by the register-fold kernel when wide_rings); GA_ERR_UNSUPPORTED (before any
// launch) if a class exceeds the instantiated sizes, and the caller then uses
// the first form
int launch_ring_synth2_f32(const ga_plan *p, int S, int G, const float *F, float *maps,
                           const ga_epilogue *epi, bool wide_rings, hipStream_t stream,
                           const float *ref = nullptr) {
  // a unit's two shells must be adjacent in the ring buffer, and the pixel
  // stores are 16-byte vectors
  if (!(G == S || G % 2 == 0)) return GA_ERR_UNSUPPORTED;
  if (reinterpret_cast<uintptr_t>(maps) & 15) return GA_ERR_UNSUPPORTED;
  if (reinterpret_cast<uintptr_t>(ref) & 15) return GA_ERR_UNSUPPORTED;
  // wide rings by the register-fold kernel (unless switched off), the other
  // power-of-two rings by the four-array kernel
  const bool wide = wide_rings && p->n_wide > 0;
  const int32_t *lp2 = wide ? p->d_rings_narrow : p->d_rings_pow2;
  const int np2 = wide ? p->n_narrow : p->n_pow2;
  // Bluestein rings: with wide_rings by the register-split kernel, the wide
  // class with the fold in registers, the rest with the fold through LDS
  // (r <= 2048, L <= 4096); otherwise all by the four-array kernel
  const bool bsplit = wide_rings && p->max_r_bnarrow <= 2048 && p->max_L_bnarrow <= 4096;
  const int32_t *lb2 = p->d_rings_blu;
  const int nb2 = bsplit ? 0 : p->n_blu;
  const int Np = wide ? p->max_r_narrow : p->max_r_pow2;
  const int Nb = p->max_L_blu, rb = p->max_r_blu;
  if (np2 > 0 && Np > 4096) return GA_ERR_UNSUPPORTED;
  if (nb2 > 0 && Nb > 4096) return GA_ERR_UNSUPPORTED;
  int rc = GA_OK;
  if (bsplit && p->n_bwide > 0) {
    rc = launch_blu_class<false>(p, p->d_rings_bwide, p->n_bwide, p->max_r_bwide, p->max_L_bwide,
                                 S, G, F, maps, epi, ref, stream);
    if (rc) return rc;
  }
  if (bsplit && p->n_bnarrow > 0) {
    rc = launch_blu_class<true>(p, p->d_rings_bnarrow, p->n_bnarrow, p->max_r_bnarrow,
                                p->max_L_bnarrow, S, G, F, maps, epi, ref, stream);
    if (rc) return rc;
  }
  if (wide) {
    switch (p->r_wide) {
      case 256: rc = launch_synth_wide<256, 1>(p, S, G, F, maps, epi, ref, stream); break;
      case 512: rc = launch_synth_wide<256, 2>(p, S, G, F, maps, epi, ref, stream); break;
      case 1024: rc = launch_synth_wide<256, 4>(p, S, G, F, maps, epi, ref, stream); break;
      case 2048: rc = launch_synth_wide<256, 8>(p, S, G, F, maps, epi, ref, stream); break;
      default: rc = launch_synth_wide<512, 8>(p, S, G, F, maps, epi, ref, stream); break;
    }
    if (rc) return rc;
  }
  if (np2 > 0) {
    const int32_t *l = lp2;
    const int n = np2;
    // LDS per workgroup: 4 arrays of N padded complex floats (67.6 KB at N =
    // 2048, 135 KB at 4096: one workgroup per CU, so 512 threads there)
    if (Np <= 256) rc = launch_synth2<256, 1, false>(p, l, n, Np, S, G, F, maps, epi, ref, stream);
    else if (Np <= 512) rc = launch_synth2<256, 2, false>(p, l, n, Np, S, G, F, maps, epi, ref, stream);
    else if (Np <= 1024) rc = launch_synth2<256, 4, false>(p, l, n, Np, S, G, F, maps, epi, ref, stream);
    else if (Np <= 2048) rc = launch_synth2<256, 8, false>(p, l, n, Np, S, G, F, maps, epi, ref, stream);
    else rc = launch_synth2<512, 8, false>(p, l, n, Np, S, G, F, maps, epi, ref, stream);
    if (rc) return rc;
  }
  if (nb2 > 0) {
    const int32_t *l = lb2;
    const int n = nb2;
    if (Nb <= 2048) {
      if (rb <= 256) rc = launch_synth2<256, 1, true>(p, l, n, Nb, S, G, F, maps, epi, ref, stream);
      else if (rb <= 512) rc = launch_synth2<256, 2, true>(p, l, n, Nb, S, G, F, maps, epi, ref, stream);
      else rc = launch_synth2<256, 4, true>(p, l, n, Nb, S, G, F, maps, epi, ref, stream);
    } else {
      rc = launch_synth2<512, 4, true>(p, l, n, Nb, S, G, F, maps, epi, ref, stream);
    }
  }
  return rc;
}

template <typename T>
int launch_ring(const ga_plan *p, int S, int G, const void *in, void *out, const ga_epilogue *epi,
                bool anal, hipStream_t stream) {
  using tab = typename tab_t<T>::type;
  const tab *blu = reinterpret_cast<const tab *>(sizeof(T) == 8 ? (const void *)p->d_blu64
                                                               : (const void *)p->d_blu32);
  const tab *tw = reinterpret_cast<const tab *>(sizeof(T) == 8 ? (const void *)p->d_tw64
                                                              : (const void *)p->d_tw32);
  const int max_r = p->max_r;
  if (anal) {
    const int npp = (S + 1) / 2;
    const int nblk = 8 * ((p->nrings + 7) / 8) * npp;
    // analysis assembles the full ring (n = 4 r <= 4 nside) in LDS as well
    const size_t lds = size_t(padded_len(std::max(p->max_L, 4 * p->nside))) * 2 * sizeof(T);
    GA_REQUIRE(lds <= 160 * 1024, GA_ERR_UNSUPPORTED,
               "ring analysis needs %zu bytes of LDS (nside %d)", lds, p->nside);
#define GA_ANAL_LAUNCH(NT_, MAXC_)                                                             \
  do {                                                                                        \
    auto kfn = ring_anal_kernel<T, NT_, MAXC_>;                                               \
    GA_CHECK_HIP(hipFuncSetAttribute((const void *)kfn,                                       \
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));  \
    hipLaunchKernelGGL(kfn, dim3(nblk), dim3(NT_), lds, stream, p->lmax, p->nrings,           \
                       p->d_ring_start, p->d_ring_n, p->d_ring_pnum, p->d_ring_pden,          \
                       p->d_blu_off, blu, tw, p->tw_len, (const T *)in, S, npp, p->npix,       \
                       T(4.0 * M_PI / double(p->npix)), (cplx<T> *)out);                       \
  } while (0)
    if (max_r <= 256) GA_ANAL_LAUNCH(256, 1);
    else if (max_r <= 1024) GA_ANAL_LAUNCH(256, 4);
    else if (max_r <= 2048) GA_ANAL_LAUNCH(256, 8);
    else if (max_r <= 4096) GA_ANAL_LAUNCH(512, 8);
    else GA_REQUIRE(false, GA_ERR_UNSUPPORTED, "ring length 4 * %d not supported", max_r);
#undef GA_ANAL_LAUNCH
    GA_LAUNCH_CHECK();
    return GA_OK;
  }
  // synthesis: P pairs per workgroup when the layout groups shells by 8 and
  // P work arrays fit in LDS
  const int ldx = padded_len(p->max_L);
  int P = 1;
  // FP32 with shells grouped by 4 or more: 2 pairs per workgroup up to
  // r = 2048 (register budget), else 1
  if (sizeof(T) == 4 && G % 4 == 0 && S > 2 && max_r <= 2048 &&
      size_t(2) * ldx * 2 * sizeof(T) <= 128 * 1024)
    P = 2;
  GA_REQUIRE(G == S || G % (2 * P) == 0, GA_ERR_INVALID, "shell group %d incompatible", G);
  const int nunit = (S + 2 * P - 1) / (2 * P);
  const int nblk = 8 * ((p->nrings + 7) / 8) * nunit;
  // P work arrays; the cooperative fold of small rings needs 4 P NT entries
  const size_t lds0 = size_t(P) * ldx * 2 * sizeof(T);
  GA_REQUIRE(lds0 <= 160 * 1024, GA_ERR_UNSUPPORTED, "ring synthesis needs %zu bytes of LDS", lds0);
#define GA_SYNTH_LAUNCH(NT_, MAXC_, P_)                                                        \
  do {                                                                                        \
    auto kfn = ring_synth_kernel<T, NT_, MAXC_, P_>;                                          \
    const size_t lds = std::max(lds0, size_t(4) * P_ * NT_ * 2 * sizeof(T));                  \
    GA_CHECK_HIP(hipFuncSetAttribute((const void *)kfn,                                       \
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));  \
    hipLaunchKernelGGL(kfn, dim3(nblk), dim3(NT_), lds, stream, p->lmax, p->nrings,           \
                       p->d_ring_start, p->d_ring_n, p->d_blu_off, blu, tw, p->tw_len, ldx,    \
                       (const cplx<T> *)in, S, G, nunit, (T *)out, p->npix, epi);             \
  } while (0)
  if constexpr (sizeof(T) == 4) {
    if (P == 2) {
      if (max_r <= 512) GA_SYNTH_LAUNCH(256, 2, 2);
      else GA_SYNTH_LAUNCH(512, 4, 2);
    }
  }
  if (P == 1) {
    if (max_r <= 256) GA_SYNTH_LAUNCH(256, 1, 1);
    else if (max_r <= 1024) GA_SYNTH_LAUNCH(256, 4, 1);
    else if (max_r <= 2048) GA_SYNTH_LAUNCH(256, 8, 1);
    else if (max_r <= 4096) GA_SYNTH_LAUNCH(512, 8, 1);
    else GA_REQUIRE(false, GA_ERR_UNSUPPORTED, "ring length 4 * %d not supported", max_r);
  }
#undef GA_SYNTH_LAUNCH
  GA_LAUNCH_CHECK();
  return GA_OK;
}

}  // namespace

int launch_ring_synth_f64(const ga_plan *p, int S, int G, const double *F, double *maps,
                          const ga_epilogue *epi, hipStream_t s) {
  return launch_ring<double>(p, S, G, F, maps, epi, false, s);
}
int launch_ring_synth_f32(const ga_plan *p, int S, int G, const float *F, float *maps,
                          const ga_epilogue *epi, hipStream_t s) {
  if (p->tune.ring_path != 1) {
    // the second form needs whole units of 2P shells in one group
    const int rc = launch_ring_synth2_f32(p, S, G, F, maps, epi, p->tune.ring_path != 2, s);
    if (rc != GA_ERR_UNSUPPORTED) return rc;
    clear_error();
  }
  return launch_ring<float>(p, S, G, F, maps, epi, false, s);
}
// out = ref - synthesis(F) in one pass (the residual of the analysis
// iterations); GA_ERR_UNSUPPORTED (nothing launched) when the fused-pass form
// is not in use or does not cover the plan, and the caller then synthesises
// and subtracts separately
int launch_ring_resid_f32(const ga_plan *p, int S, int G, const float *F, float *out,
                          const float *ref, hipStream_t s) {
  if (p->tune.ring_path == 1) return GA_ERR_UNSUPPORTED;
  const int rc = launch_ring_synth2_f32(p, S, G, F, out, nullptr, p->tune.ring_path != 2, s, ref);
  if (rc == GA_ERR_UNSUPPORTED) clear_error();
  return rc;
}
int launch_ring_anal_f64(const ga_plan *p, int S, const double *maps, double *F, hipStream_t s) {
  return launch_ring<double>(p, S, S, maps, F, nullptr, true, s);
}
int launch_ring_anal_f32(const ga_plan *p, int S, const float *maps, float *F, hipStream_t s) {
  if (p->tune.ring_path != 1) {
    const int rc = launch_ring_anal2_f32(p, S, maps, F, s);
    if (rc != GA_ERR_UNSUPPORTED) return rc;
    clear_error();
  }
  return launch_ring<float>(p, S, S, maps, F, nullptr, true, s);
}

}  // namespace ga
