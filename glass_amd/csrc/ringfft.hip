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

// In-place forward FFT (e^{-2 pi i}), natural order in, bit-reversed out.
template <typename T, int NT>
__device__ void fft_dif(cplx<T> *X, int L, const typename tab_t<T>::type *tw, int maxL) {
  const int t = threadIdx.x;
  const int logL = ilog2(L);
  int h = L >> 1;
  if (logL & 1) {
    const int stride = maxL / (2 * h);
    for (int j = t; j < h; j += NT) {
      cplx<T> w = cconj(ld<T>(tw, int64_t(j) * stride));
      cplx<T> u = X[j], v = X[j + h];
      X[j] = cadd(u, v);
      X[j + h] = cmul(csub(u, v), w);
    }
    __syncthreads();
    h >>= 1;
  }
  for (; h >= 2; h >>= 2) {
    const int q = h >> 1;
    const int stride = maxL / (2 * h);
    for (int g = t; g < (L >> 2); g += NT) {
      const int blk = g / q, pos = g - blk * q;
      const int b0 = blk * 2 * h + pos;
      cplx<T> w1 = cconj(ld<T>(tw, int64_t(pos) * stride));
      cplx<T> w2 = cmul(w1, w1);
      cplx<T> w3 = {w1.im, -w1.re};  // w1 * (-i)
      cplx<T> a = X[b0], b = X[b0 + q], c = X[b0 + h], d = X[b0 + h + q];
      cplx<T> a1 = cadd(a, c), c1 = cmul(csub(a, c), w1);
      cplx<T> b1 = cadd(b, d), d1 = cmul(csub(b, d), w3);
      X[b0] = cadd(a1, b1);
      X[b0 + q] = cmul(csub(a1, b1), w2);
      X[b0 + h] = cadd(c1, d1);
      X[b0 + h + q] = cmul(csub(c1, d1), w2);
    }
    __syncthreads();
  }
}

// In-place inverse FFT (e^{+2 pi i}, unscaled), bit-reversed in, natural out.
template <typename T, int NT>
__device__ void fft_dit(cplx<T> *X, int L, const typename tab_t<T>::type *tw, int maxL) {
  const int t = threadIdx.x;
  const int logL = ilog2(L);
  int h = 1;
  if (logL & 1) {
    for (int j = t; j < (L >> 1); j += NT) {
      cplx<T> u = X[2 * j], v = X[2 * j + 1];
      X[2 * j] = cadd(u, v);
      X[2 * j + 1] = csub(u, v);
    }
    __syncthreads();
    h = 2;
  }
  for (; h < L; h <<= 2) {
    const int stride = maxL / (4 * h);
    for (int g = t; g < (L >> 2); g += NT) {
      const int blk = g / h, pos = g - blk * h;
      const int a0 = blk * 4 * h + pos;
      cplx<T> w4 = ld<T>(tw, int64_t(pos) * stride);
      cplx<T> w = cmul(w4, w4);
      cplx<T> w4i = cmuli(w4);
      cplx<T> a = X[a0], b = X[a0 + h], c = X[a0 + 2 * h], d = X[a0 + 3 * h];
      cplx<T> tb = cmul(b, w), td = cmul(d, w);
      cplx<T> a1 = cadd(a, tb), b1 = csub(a, tb);
      cplx<T> c1 = cadd(c, td), d1 = csub(c, td);
      cplx<T> tc = cmul(c1, w4), tdd = cmul(d1, w4i);
      X[a0] = cadd(a1, tc);
      X[a0 + 2 * h] = csub(a1, tc);
      X[a0 + h] = cadd(b1, tdd);
      X[a0 + 3 * h] = csub(b1, tdd);
    }
    __syncthreads();
  }
}

__device__ __forceinline__ int blu_len(int r) {
  int L = 1;
  while (L < 2 * r - 1) L <<= 1;
  return L;
}

// Inverse length-r DFT of the values H[i][slot] (owned j1 = t + i NT), in place
// in the registers; X is the LDS work array.  conj_io conjugates input and
// output, which turns it into the forward DFT.
template <typename T, int NT, int MAXC>
__device__ void sub_idft(cplx<T> (&H)[MAXC][4], int slot, int r, cplx<T> *X, int64_t boff,
                         const typename tab_t<T>::type *blu, const typename tab_t<T>::type *tw,
                         int maxL, bool conj_io) {
  const int t = threadIdx.x;
  __syncthreads();  // X free
  if (boff < 0) {   // r is a power of two
    const int bits = ilog2(r);
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      const int j1 = t + i * NT;
      if (j1 < r) {
        cplx<T> v = H[i][slot];
        if (conj_io) v = cconj(v);
        const int pos = bits ? int(__brev(unsigned(j1)) >> (32 - bits)) : 0;
        X[pos] = v;
      }
    }
    __syncthreads();
    fft_dit<T, NT>(X, r, tw, maxL);
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      const int j1 = t + i * NT;
      if (j1 < r) {
        cplx<T> v = X[j1];
        H[i][slot] = conj_io ? cconj(v) : v;
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
        cplx<T> v = H[i][slot];
        if (conj_io) v = cconj(v);
        X[j1] = cmul(v, ld<T>(c, j1));
      }
    }
    for (int j = r + t; j < L; j += NT) X[j] = cplx<T>{T(0), T(0)};
    __syncthreads();
    fft_dif<T, NT>(X, L, tw, maxL);
    for (int j = t; j < L; j += NT) X[j] = cmul(X[j], ld<T>(bh, j));
    __syncthreads();
    fft_dit<T, NT>(X, L, tw, maxL);
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      const int j1 = t + i * NT;
      if (j1 < r) {
        cplx<T> v = cmul(X[j1], ld<T>(c, j1));
        H[i][slot] = conj_io ? cconj(v) : v;
      }
    }
  }
}

template <typename T>
__device__ __forceinline__ T apply_epi(T x, int kind, T lam, T par) {
  if (kind == GA_T_LOGNORMAL) {
    T y = expm1(x - par * T(0.5));
    return lam == T(1) ? y : lam * y;
  }
  if (kind == GA_T_SQUARED_NORMAL) {
    T d = x - par;
    T y = d * d - T(1);
    return lam == T(1) ? y : lam * y;
  }
  return x;
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

template <typename T, int NT, int MAXC>
__global__ __launch_bounds__(NT) void ring_synth_kernel(
    int lmax, int nrings, const int64_t *__restrict__ ring_start,
    const int32_t *__restrict__ ring_n, const int32_t *__restrict__ ring_pnum,
    const int32_t *__restrict__ ring_pden, const int64_t *__restrict__ blu_off,
    const typename tab_t<T>::type *__restrict__ blu, const typename tab_t<T>::type *__restrict__ tw,
    int maxL, const cplx<T> *__restrict__ F, int S, int npp, T *__restrict__ maps, int64_t npix,
    const ga_epilogue *__restrict__ epi) {
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
  const int64_t frow = int64_t(ring) * (lmax + 1);

  cplx<T> H[MAXC][4];
  // ---- fold coefficients into the n bins, bins k = 4 j1 + k2 owned by this thread
#pragma unroll
  for (int i = 0; i < MAXC; ++i) {
    const int j1 = t + i * NT;
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
      cplx<T> g0 = {T(0), T(0)}, g1 = {T(0), T(0)};
      if (j1 < r) {
        const int k = 4 * j1 + k2;
        for (int m = k; m <= lmax; m += n) {
          const cplx<T> ph = pnum ? expi_q(m % pden, pden, T(0)) : cplx<T>{T(1), T(0)};
          const cplx<T> *f = F + (frow + m) * S;
          g0 = cadd(g0, cmul(f[s0], ph));
          if (has1) g1 = cadd(g1, cmul(f[s1], ph));
        }
        for (int m = (k == 0 ? n : n - k); m <= lmax; m += n) {
          const cplx<T> ph = pnum ? expi_q(m % pden, pden, T(0)) : cplx<T>{T(1), T(0)};
          const cplx<T> *f = F + (frow + m) * S;
          g0 = cadd(g0, cconj(cmul(f[s0], ph)));
          if (has1) g1 = cadd(g1, cconj(cmul(f[s1], ph)));
        }
      }
      H[i][k2] = cplx<T>{g0.re - g1.im, g0.im + g1.re};
    }
  }
  // ---- four length-r inverse DFTs
#pragma unroll
  for (int k2 = 0; k2 < 4; ++k2) sub_idft<T, NT, MAXC>(H, k2, r, X, boff, blu, tw, maxL, false);

  // ---- radix-4 combine, epilogue, store
  const int64_t st = ring_start[ring];
  T lam0 = T(1), par0 = T(0), lam1 = T(1), par1 = T(0);
  int kind0 = GA_T_NORMAL, kind1 = GA_T_NORMAL;
  if (epi) {
    kind0 = epi[s0].kind; lam0 = T(epi[s0].lamda); par0 = T(epi[s0].param);
    if (has1) { kind1 = epi[s1].kind; lam1 = T(epi[s1].lamda); par1 = T(epi[s1].param); }
  }
  T *out0 = maps + int64_t(s0) * npix + st;
  T *out1 = maps + int64_t(s1) * npix + st;
#pragma unroll
  for (int i = 0; i < MAXC; ++i) {
    const int j1 = t + i * NT;
    if (j1 < r) {
      const cplx<T> y0 = H[i][0];
      const cplx<T> y1 = cmul(H[i][1], expi_q(j1, n, T(0)));
      const cplx<T> y2 = cmul(H[i][2], expi_q((2 * j1) % n, n, T(0)));
      const cplx<T> y3 = cmul(H[i][3], expi_q((3 * j1) % n, n, T(0)));
      const cplx<T> s02 = cadd(y0, y2), d02 = csub(y0, y2);
      const cplx<T> s13 = cadd(y1, y3), d13 = cmuli(csub(y1, y3));
      const cplx<T> o[4] = {cadd(s02, s13), cadd(d02, d13), csub(s02, s13), csub(d02, d13)};
#pragma unroll
      for (int j2 = 0; j2 < 4; ++j2) {
        const int64_t pix = j1 + int64_t(r) * j2;
        out0[pix] = apply_epi<T>(o[j2].re, kind0, lam0, par0);
        if (has1) out1[pix] = apply_epi<T>(o[j2].im, kind1, lam1, par1);
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
  cplx<T> H[MAXC][4];
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
      H[i][j2] = v;
    }
  }
  // forward length-r DFTs of each decimated sequence
#pragma unroll
  for (int j2 = 0; j2 < 4; ++j2) sub_idft<T, NT, MAXC>(H, j2, r, X, boff, blu, tw, maxL, true);
  __syncthreads();
  // X_{k1 + r k2} = sum_{j2} (-i)^{j2 k2} e^{-2 pi i j2 k1 / n} D_{j2}[k1] -> LDS
#pragma unroll
  for (int i = 0; i < MAXC; ++i) {
    const int k1 = t + i * NT;
    if (k1 < r) {
      const cplx<T> y0 = H[i][0];
      const cplx<T> y1 = cmul(H[i][1], cconj(expi_q(k1, n, T(0))));
      const cplx<T> y2 = cmul(H[i][2], cconj(expi_q((2 * k1) % n, n, T(0))));
      const cplx<T> y3 = cmul(H[i][3], cconj(expi_q((3 * k1) % n, n, T(0))));
      const cplx<T> s02 = cadd(y0, y2), d02 = csub(y0, y2);
      const cplx<T> s13 = cadd(y1, y3), d13 = cmuli(csub(y3, y1));  // -i (y1 - y3)
      X[k1] = cadd(s02, s13);
      X[k1 + r] = cadd(d02, d13);
      X[k1 + 2 * r] = csub(s02, s13);
      X[k1 + 3 * r] = csub(d02, d13);
    }
  }
  __syncthreads();
  // split the two real transforms and unfold into m = 0..lmax
  for (int m = t; m <= lmax; m += NT) {
    const int k = m % n;
    const int kk = k ? n - k : 0;
    const cplx<T> a = X[k], b = cconj(X[kk]);
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

template <typename T>
int launch_ring(const ga_plan *p, int S, const void *in, void *out, const ga_epilogue *epi,
                bool anal, hipStream_t stream) {
  const int npp = (S + 1) / 2;
  const int nblk = 8 * ((p->nrings + 7) / 8) * npp;
  // synthesis needs one length-L work array; analysis also assembles the
  // full ring (n = 4 r <= 4 nside) there
  const int lds_elems = anal ? std::max(p->max_L, 4 * p->nside) : p->max_L;
  const size_t lds = size_t(lds_elems) * 2 * sizeof(T);
  GA_REQUIRE(lds <= 160 * 1024, GA_ERR_UNSUPPORTED,
             "ring transform needs %zu bytes of LDS (nside %d, %zu-byte reals)", lds, p->nside,
             sizeof(T));
  using tab = typename tab_t<T>::type;
  const tab *blu = reinterpret_cast<const tab *>(sizeof(T) == 8 ? (const void *)p->d_blu64
                                                                : (const void *)p->d_blu32);
  const tab *tw = reinterpret_cast<const tab *>(sizeof(T) == 8 ? (const void *)p->d_tw64
                                                               : (const void *)p->d_tw32);
#define GA_RING_LAUNCH(NT_, MAXC_)                                                             \
  do {                                                                                        \
    if (anal) {                                                                               \
      auto kfn = ring_anal_kernel<T, NT_, MAXC_>;                                             \
      GA_CHECK_HIP(hipFuncSetAttribute((const void *)kfn,                                     \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));\
      hipLaunchKernelGGL(kfn, dim3(nblk), dim3(NT_), lds, stream, p->lmax, p->nrings,         \
                         p->d_ring_start, p->d_ring_n, p->d_ring_pnum, p->d_ring_pden,        \
                         p->d_blu_off, blu, tw, p->max_L, (const T *)in, S, npp, p->npix,     \
                         T(4.0 * M_PI / double(p->npix)), (cplx<T> *)out);                     \
    } else {                                                                                  \
      auto kfn = ring_synth_kernel<T, NT_, MAXC_>;                                            \
      GA_CHECK_HIP(hipFuncSetAttribute((const void *)kfn,                                     \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));\
      hipLaunchKernelGGL(kfn, dim3(nblk), dim3(NT_), lds, stream, p->lmax, p->nrings,         \
                         p->d_ring_start, p->d_ring_n, p->d_ring_pnum, p->d_ring_pden,        \
                         p->d_blu_off, blu, tw, p->max_L, (const cplx<T> *)in, S, npp,        \
                         (T *)out, p->npix, epi);                                             \
    }                                                                                         \
  } while (0)
  const int max_r = p->max_r;
  if (max_r <= 256) GA_RING_LAUNCH(256, 1);
  else if (max_r <= 1024) GA_RING_LAUNCH(256, 4);
  else if (max_r <= 2048) GA_RING_LAUNCH(256, 8);
  else if (max_r <= 4096) GA_RING_LAUNCH(512, 8);
  else GA_REQUIRE(false, GA_ERR_UNSUPPORTED, "ring length 4 * %d not supported", max_r);
#undef GA_RING_LAUNCH
  GA_LAUNCH_CHECK();
  return GA_OK;
}

}  // namespace

int launch_ring_synth_f64(const ga_plan *p, int S, const double *F, double *maps,
                          const ga_epilogue *epi, hipStream_t s) {
  return launch_ring<double>(p, S, F, maps, epi, false, s);
}
int launch_ring_synth_f32(const ga_plan *p, int S, const float *F, float *maps,
                          const ga_epilogue *epi, hipStream_t s) {
  return launch_ring<float>(p, S, F, maps, epi, false, s);
}
int launch_ring_anal_f64(const ga_plan *p, int S, const double *maps, double *F, hipStream_t s) {
  return launch_ring<double>(p, S, maps, F, nullptr, true, s);
}
int launch_ring_anal_f32(const ga_plan *p, int S, const float *maps, float *F, hipStream_t s) {
  return launch_ring<float>(p, S, maps, F, nullptr, true, s);
}

}  // namespace ga
