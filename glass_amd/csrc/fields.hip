// Field-side kernels of the generation path: the conditional-Gaussian scaling
// vectors (iternorm), the correlated alm of each shell (multalm sum, GLASS ->
// healpy reorder and the m = 0 fix fused into one pass), almxfl, and the
// pointwise transforms.  All of it is HBM-bound elementwise work.
#include <cmath>

#include "common.h"
#include "device.h"
#include "kernels.h"

namespace ga {

namespace {

// healpy index i -> (l, m) for mmax = lmax = L:  i = m (2L + 1 - m) / 2 + l
__device__ __forceinline__ void hp_decode(int64_t i, int L, int &l, int &m) {
  const double b = 2.0 * L + 1.0;
  int mm = int((b - sqrt(b * b - 8.0 * double(i))) * 0.5);
  if (mm < 0) mm = 0;
  while (mm > 0 && hp_index(L, mm, mm) > i) --mm;
  while (mm < L && hp_index(L, mm + 1, mm + 1) <= i) ++mm;
  m = mm;
  l = int(i - hp_index(L, mm, mm)) + mm;
}

// GLASS index g = l (l + 1) / 2 + m -> l
__device__ __forceinline__ int glass_l(int64_t g) {
  int l = int((sqrt(8.0 * double(g) + 1.0) - 1.0) * 0.5);
  while (int64_t(l) * (l + 1) / 2 > g) --l;
  while (int64_t(l + 1) * (l + 2) / 2 <= g) ++l;
  return l;
}

// Standard complex normal pairs (re, im) for GLASS-order index g: one
// Philox4x32-10 call with counter (g lo, g hi, u lo, u hi) serves the shells
// t = 2u (words 0, 1) and t = 2u + 1 (words 2, 3).  Box-Muller on a word pair
// (w0, w1): u1 = (w0 + 1/2) 2^-32 in (0, 1] (tail to sqrt(-2 ln 2^-33) =
// 6.76 sigma) and u2 = (w1 >> 8) 2^-24 in [0, 1), with the hardware log2 /
// sqrt / sin(2 pi x) / cos(2 pi x) in FP32 (relative error a few 1e-7: far
// below what the statistics of the field can resolve), widened to FP64 so
// that every path (FP32 batch, FP64 reference) sees the same values.
__device__ __forceinline__ u32x4 device_normal_words(uint64_t seed, int64_t u, int64_t g) {
  u32x4 c{{uint32_t(g), uint32_t(uint64_t(g) >> 32), uint32_t(u), uint32_t(uint64_t(u) >> 32)}};
  return philox4x32_10(c, uint32_t(seed), uint32_t(seed >> 32));
}
__device__ __forceinline__ double2 normal_from_words(uint32_t w0, uint32_t w1) {
  const float u1 = (float(w0) + 0.5f) * 0x1p-32f;
  const float u2 = float(w1 >> 8) * 0x1p-24f;
  const float rad = __builtin_sqrtf(-2.0f * 0.69314718055994531f * __builtin_amdgcn_logf(u1));
  const float sn = __builtin_amdgcn_sinf(u2), cs = __builtin_amdgcn_cosf(u2);
  return double2{double(rad * cs), double(rad * sn)};
}
// the pair of shell t (t >= 0)
__device__ __forceinline__ double2 device_normal_pair(uint64_t seed, int64_t t, int64_t g) {
  const u32x4 r = device_normal_words(seed, t >> 1, g);
  return (t & 1) ? normal_from_words(r.v[2], r.v[3]) : normal_from_words(r.v[0], r.v[1]);
}

// State per mode l (capacity kmax): M [kmax][kmax] row-major, a [kmax], s,
// then kmax doubles of scratch.
// Row t of this call has k + 1 entries; the state entering it has size kprev
// (= k for every row after the first of a call).  Update (arXiv:2302.01942
// App. A): grow M to (kprev+1)^2 with the new bottom row [-a^T M, u] / s'
// (u = 1, s' = s where s > 0; u = 0, s' = 1 otherwise), keep the trailing
// k x k block, then a = M c (c = correlations oldest first), s^2 = C_0 - a.a.
__global__ void iternorm_kernel(int nl, int kprev0, int k, int kmax, int nshell, int start,
                                const double *__restrict__ cov, double *__restrict__ state,
                                double *__restrict__ w, int *__restrict__ flag) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nl) return;
  double *M = state + int64_t(l) * (kmax * kmax + 2 * kmax + 1);
  double *a = M + kmax * kmax;
  double *s = a + kmax;
  double *atm = s + 1;
  int kprev = kprev0;
  for (int t = 0; t < nshell; ++t) {
    const double *row = cov + (int64_t(t) * nl + l) * (k + 1);
    double *wt = w + (int64_t(t) * nl + l) * (k + 1);
    if (t == 0 && start) {
      for (int i = 0; i < k; ++i) {
        for (int j = 0; j < k; ++j) M[i * kmax + j] = 0.0;
        a[i] = 0.0;
      }
      *s = 1.0;
    } else {
      const double sv = *s;
      const double u = sv > 0.0 ? 1.0 : 0.0;
      const double sp = sv > 0.0 ? sv : 1.0;
      // atm <- a^T M / s'
      for (int j = 0; j < kprev; ++j) {
        double acc = 0.0;
        for (int i = 0; i < kprev; ++i) acc += a[i] * M[i * kmax + j];
        atm[j] = acc / sp;
      }
      const int d = kprev + 1 - k;  // rows/cols dropped from the grown matrix
      for (int i = 0; i < k; ++i) {
        const int bi = i + d;
        for (int j = 0; j < k; ++j) {
          const int bj = j + d;
          double v;
          if (bi < kprev) v = (bj < kprev) ? M[bi * kmax + bj] : 0.0;
          else v = (bj < kprev) ? -atm[bj] : u / sp;
          M[i * kmax + j] = v;
        }
      }
    }
    double aa = 0.0;
    for (int i = 0; i < k; ++i) {
      double acc = 0.0;
      for (int j = 0; j < k; ++j) acc += M[i * kmax + j] * row[k - j];
      a[i] = acc;
      aa += acc * acc;
    }
    const double var = row[0] - aa;
    if (var < 0.0) atomicOr(flag, 1);
    *s = sqrt(var > 0.0 ? var : 0.0);
    for (int i = 0; i < k; ++i) wt[i] = a[i];
    wt[k] = *s;
    kprev = k;
  }
}

// One thread per (healpy index, shell); shell index innermost for coalescing.
template <bool HOSTZ>
__global__ void fields_alm_kernel(int nl, int k, int first_shell, int nshell, int64_t nalm,
                                  const double *__restrict__ w, const double2 *__restrict__ zring,
                                  int nslot, uint64_t seed, float2 *__restrict__ out32, int ldo,
                                  double2 *__restrict__ out64) {
  const int64_t gid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= nalm * nshell) return;
  const int64_t i = gid / nshell;
  const int s = int(gid - i * nshell);
  int l, m;
  hp_decode(i, nl - 1, l, m);
  const int64_t g = int64_t(l) * (l + 1) / 2 + m;
  const int64_t z_size = int64_t(nl) * (nl + 1) / 2;
  const int64_t t = int64_t(first_shell) + s;
  const int len = int(t + 1 < k + 1 ? t + 1 : k + 1);
  const int mis = k + 1 - len;
  const double *wl = w + (int64_t(s) * nl + l) * (k + 1);
  double re = 0.0, im = 0.0;
  for (int j = mis; j <= k; ++j) {
    const int64_t tz = t - k + j;  // shell whose normals enter with weight w[j]
    double2 z;
    if (HOSTZ)
      z = zring[(tz % nslot) * z_size + g];
    else
      z = device_normal_pair(seed, tz, g);
    re = fma(wl[j], z.x, re);
    im = fma(wl[j], z.y, im);
  }
  if (m == 0) { re += im; im = 0.0; }
  if (out32) out32[i * ldo + s] = float2{float(re), float(im)};
  else out64[i] = double2{re, im};
}

// Device-RNG form, one thread per healpy index walking the shells in order:
// each normal vector entry z_t[g] is generated once and kept in a sliding
// register window of K+1 values (K, the number of correlated neighbours, is a
// template argument so that the window moves by register renaming rather than
// by per-entry selects); the weights of the next shell are loaded while the
// current one is generated; the [idx][S] output is staged through LDS in
// chunks of 16 shells so that stores are contiguous 128-byte rows.  With
// colpart, block b also writes the maxima of |re| and |im| of its rows per
// shell (colpart[b][2 s + 0/1], the float bits ordered as unsigned integers,
// as the operand preparation orders them), so that the column maxima of the
// whole batch take one pass over ceil(nalm / 256) rows instead of nalm.
constexpr int kWinMax = 8;  // K + 1 <= kWinMax
constexpr int kFaBlock = 256, kFaChunk = 16;
static_assert(kFaBlock % 64 == 0 && kFaBlock % kFaChunk == 0 && 2 * kFaChunk <= 32,
              "column-maximum reduction layout");

template <int K>
__global__ __launch_bounds__(kFaBlock) void fields_alm_window_kernel(
    int nl, int first_shell, int nshell, int64_t nalm, const double *__restrict__ w,
    uint64_t seed, float2 *__restrict__ out32, int ldo, float *__restrict__ colpart) {
  __shared__ float2 tile[kFaBlock][kFaChunk + 1];
  __shared__ uint32_t red[kFaBlock / 32][32];
  const int tid = threadIdx.x;
  const int64_t i0 = int64_t(blockIdx.x) * kFaBlock;
  const int64_t i = i0 + tid;
  const bool valid = i < nalm;
  int l = 0, m = 0;
  if (valid) hp_decode(i, nl - 1, l, m);
  const int64_t g = int64_t(l) * (l + 1) / 2 + m;
  // window zz[j] = z_{t - K + j} for the current shell t; (sw0, sw1): the
  // Philox words of shell t + 1 when t is even (one call serves both)
  double2 zz[K + 1];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int64_t tz = int64_t(first_shell) - K + j;
    zz[j] = (tz >= 0 && valid) ? device_normal_pair(seed, tz, g) : double2{0.0, 0.0};
  }
  uint32_t sw0, sw1;
  {
    const u32x4 r = device_normal_words(seed, int64_t(first_shell) >> 1, g);
    const bool odd = first_shell & 1;
    zz[K] = valid ? normal_from_words(odd ? r.v[2] : r.v[0], odd ? r.v[3] : r.v[1])
                  : double2{0.0, 0.0};
    sw0 = r.v[2];
    sw1 = r.v[3];
  }
  // weights of shell s for this l: wrow[s * wstride + j], j = 0..K
  const double *wrow = w + int64_t(l) * (K + 1);
  const int64_t wstride = int64_t(nl) * (K + 1);
  double wc[K + 1];
#pragma unroll
  for (int j = 0; j <= K; ++j) wc[j] = wrow[j];
  const int nrows = int(nalm - i0 < kFaBlock ? nalm - i0 : kFaBlock);
  for (int s0 = 0; s0 < nshell; s0 += kFaChunk) {
    const int cnt = nshell - s0 < kFaChunk ? nshell - s0 : kFaChunk;
    for (int c = 0; c < cnt; ++c) {
      const int s = s0 + c;
      const int64_t t = int64_t(first_shell) + s;
      const int sn = s + 1 < nshell ? s + 1 : s;
      double wn[K + 1];
#pragma unroll
      for (int j = 0; j <= K; ++j) wn[j] = wrow[sn * wstride + j];
      if (s > 0) {
#pragma unroll
        for (int j = 0; j < K; ++j) zz[j] = zz[j + 1];
        // t odd: shell t - 1 (the previous one) was even and left its
        // partner's words; t even: a new call (block-uniform branch)
        uint32_t w0 = sw0, w1 = sw1;
        if (!(t & 1)) {
          const u32x4 r = device_normal_words(seed, t >> 1, g);
          w0 = r.v[0];
          w1 = r.v[1];
          sw0 = r.v[2];
          sw1 = r.v[3];
        }
        zz[K] = valid ? normal_from_words(w0, w1) : double2{0.0, 0.0};
      }
      double re = 0.0, im = 0.0;
      if (t >= K) {
#pragma unroll
        for (int j = 0; j <= K; ++j) {
          re = fma(wc[j], zz[j].x, re);
          im = fma(wc[j], zz[j].y, im);
        }
      } else {
        // rows of the first K shells have t + 1 < K + 1 entries: skip the leading ones
        const int mis = int(K - t);
#pragma unroll
        for (int j = 0; j <= K; ++j) {
          if (j >= mis) {
            re = fma(wc[j], zz[j].x, re);
            im = fma(wc[j], zz[j].y, im);
          }
        }
      }
      if (m == 0) { re += im; im = 0.0; }
      tile[tid][c] = float2{float(re), float(im)};
#pragma unroll
      for (int j = 0; j <= K; ++j) wc[j] = wn[j];
    }
    __syncthreads();
    // cooperative store: element (r, c) of the tile -> out[(i0 + r) ldo + s0 + c]
    uint32_t mre = 0u, mim = 0u;  // maxima of what this thread stored
    for (int e = tid; e < nrows * cnt; e += kFaBlock) {
      const int r = e / cnt, c = e - r * cnt;
      const float2 v = tile[r][c];
      out32[(i0 + r) * ldo + s0 + c] = v;
      mre = max(mre, __float_as_uint(fabsf(v.x)));
      mim = max(mim, __float_as_uint(fabsf(v.y)));
    }
    if (colpart && cnt == kFaChunk) {
      // a full chunk: thread tid stored column tid % 16 only, so lanes
      // tid ^ 16, ^ 32, ^ 48 of its wave and then the four waves combine
      const int lane = tid & 63, wave = tid >> 6;
      mre = max(mre, uint32_t(__shfl_xor(int(mre), 16, 64)));
      mim = max(mim, uint32_t(__shfl_xor(int(mim), 16, 64)));
      mre = max(mre, uint32_t(__shfl_xor(int(mre), 32, 64)));
      mim = max(mim, uint32_t(__shfl_xor(int(mim), 32, 64)));
      if (lane < kFaChunk) {
        red[wave][2 * lane] = mre;
        red[wave][2 * lane + 1] = mim;
      }
      __syncthreads();
      if (tid < 2 * kFaChunk) {
        uint32_t m2 = red[0][tid];
#pragma unroll
        for (int g = 1; g < kFaBlock / 64; ++g) m2 = max(m2, red[g][tid]);
        colpart[int64_t(blockIdx.x) * (2 * nshell) + 2 * s0 + tid] = __uint_as_float(m2);
      }
    } else if (colpart) {
      // a partial chunk: thread (rg, cc) takes component cc (re / im of
      // shell s0 + cc / 2) over rows rg, rg + 8, ...
      const int cc = tid & 31, rg = tid >> 5;
      uint32_t mx = 0u;
      if (cc < 2 * cnt)
        for (int r = rg; r < nrows; r += kFaBlock / 32) {
          const float2 v = tile[r][cc >> 1];
          mx = max(mx, __float_as_uint(fabsf((cc & 1) ? v.y : v.x)));
        }
      red[rg][cc] = mx;
      __syncthreads();
      if (tid < 2 * cnt) {
        uint32_t m2 = red[0][tid];
#pragma unroll
        for (int g = 1; g < kFaBlock / 32; ++g) m2 = max(m2, red[g][tid]);
        colpart[int64_t(blockIdx.x) * (2 * nshell) + 2 * s0 + tid] = __uint_as_float(m2);
      }
    }
    __syncthreads();
  }
}

template <int K>
void launch_fields_window(int nl, int first_shell, int nshell, int64_t nalm, const double *w,
                          uint64_t seed, float2 *out32, int ldo, float *colpart,
                          hipStream_t stream) {
  hipLaunchKernelGGL(fields_alm_window_kernel<K>, dim3(unsigned((nalm + kFaBlock - 1) / kFaBlock)),
                     dim3(kFaBlock), 0, stream, nl, first_shell, nshell, nalm, w, seed, out32, ldo,
                     colpart);
}

__global__ void multalm_kernel(int64_t n, const double2 *__restrict__ alm,
                               const double *__restrict__ bl, double2 *__restrict__ out) {
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const double f = bl[glass_l(g)];
  const double2 a = alm[g];
  out[g] = double2{a.x * f, a.y * f};
}

__global__ void glass_to_healpix_kernel(int nl, int64_t n, const double2 *__restrict__ in,
                                        double2 *__restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int l, m;
  hp_decode(i, nl - 1, l, m);
  out[i] = in[int64_t(l) * (l + 1) / 2 + m];
}

template <typename T>
__global__ void almxfl_kernel(int lmax, int nshell, int64_t n, const T *__restrict__ alm,
                              const double *__restrict__ fl, int nfl, T *__restrict__ out) {
  const int64_t gid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= n) return;
  const int64_t i = gid / nshell;
  int l, m;
  hp_decode(i, lmax, l, m);
  const double f = l < nfl ? fl[l] : 0.0;
  const T a = alm[gid];
  out[gid] = T{decltype(a.x)(a.x * f), decltype(a.y)(a.y * f)};
}

__global__ void transform_f64_kernel(int64_t n, const double *__restrict__ x, int kind,
                                     double lam, double par, double *__restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = x[i];
  if (kind == GA_T_LOGNORMAL) {
    v = expm1(v - par * 0.5);
    if (lam != 1.0) v *= lam;
  } else if (kind == GA_T_SQUARED_NORMAL) {
    v = (v - par) * (v - par) - 1.0;
    if (lam != 1.0) v *= lam;
  }
  out[i] = v;
}

inline unsigned nblocks(int64_t n, int bs) { return unsigned((n + bs - 1) / bs); }
void launch_window(int nl, int k, int first_shell, int nshell, int64_t nalm, const double *w,
                   uint64_t seed, float2 *o, int ldo, float *colpart, hipStream_t st) {
  switch (k) {
    case 0: launch_fields_window<0>(nl, first_shell, nshell, nalm, w, seed, o, ldo, colpart, st); break;
    case 1: launch_fields_window<1>(nl, first_shell, nshell, nalm, w, seed, o, ldo, colpart, st); break;
    case 2: launch_fields_window<2>(nl, first_shell, nshell, nalm, w, seed, o, ldo, colpart, st); break;
    case 3: launch_fields_window<3>(nl, first_shell, nshell, nalm, w, seed, o, ldo, colpart, st); break;
    case 4: launch_fields_window<4>(nl, first_shell, nshell, nalm, w, seed, o, ldo, colpart, st); break;
    case 5: launch_fields_window<5>(nl, first_shell, nshell, nalm, w, seed, o, ldo, colpart, st); break;
    case 6: launch_fields_window<6>(nl, first_shell, nshell, nalm, w, seed, o, ldo, colpart, st); break;
    default: launch_fields_window<7>(nl, first_shell, nshell, nalm, w, seed, o, ldo, colpart, st); break;
  }
}


}  // namespace

}  // namespace ga

using namespace ga;

extern "C" int ga_iternorm_state_doubles(int32_t kmax, int32_t *n) {
  GA_REQUIRE(n && kmax >= 0, GA_ERR_INVALID, "invalid argument");
  *n = kmax * kmax + 2 * kmax + 1;
  return GA_OK;
}

extern "C" int ga_iternorm_batch(int32_t nl, int32_t kprev, int32_t k, int32_t kmax,
                                 int32_t nshell, int32_t start, const double *cov, double *state,
                                 double *w, int32_t *flag, void *stream) {
  clear_error();
  GA_REQUIRE(nl > 0 && k >= 0 && kmax >= k && nshell > 0 && cov && state && w && flag,
             GA_ERR_INVALID, "invalid argument to ga_iternorm_batch");
  GA_REQUIRE(start || (kprev >= 0 && kprev <= kmax && k <= kprev + 1), GA_ERR_INVALID,
             "shape mismatch in covariance (%d correlations after %d)", k, kprev);
  hipLaunchKernelGGL(iternorm_kernel, dim3(nblocks(nl, 128)), dim3(128), 0, (hipStream_t)stream,
                     nl, kprev, k, kmax, nshell, start, cov, state, w, flag);
  GA_LAUNCH_CHECK();
  return GA_OK;
}

extern "C" int ga_fields_alm(int32_t nl, int32_t k, int32_t first_shell, int32_t nshell,
                             const double *w, const double *zring, int32_t nslot, uint64_t seed,
                             float *out_f32, int32_t ldo, double *out_f64, void *stream) {
  clear_error();
  GA_REQUIRE(nl > 0 && k >= 0 && first_shell >= 0 && nshell > 0 && w, GA_ERR_INVALID,
             "invalid argument to ga_fields_alm");
  GA_REQUIRE((out_f32 != nullptr) != (out_f64 != nullptr), GA_ERR_INVALID,
             "exactly one of out_f32 / out_f64 must be given");
  GA_REQUIRE(out_f64 == nullptr || nshell == 1, GA_ERR_INVALID, "out_f64 requires nshell == 1");
  if (ldo == 0) ldo = nshell;
  GA_REQUIRE(ldo >= nshell, GA_ERR_INVALID, "ldo %d < nshell %d", ldo, nshell);
  GA_REQUIRE(zring == nullptr || nslot >= nshell + k, GA_ERR_INVALID,
             "normal ring buffer needs nslot >= nshell + k (%d < %d)", nslot, nshell + k);
  const int64_t nalm = int64_t(nl) * (nl + 1) / 2;
  const int64_t n = nalm * nshell;
  const dim3 grid(nblocks(n, 256)), block(256);
  if (!zring && out_f32 && k + 1 <= kWinMax) {
    launch_window(nl, k, first_shell, nshell, nalm, w, seed, reinterpret_cast<float2 *>(out_f32),
                  ldo, nullptr, (hipStream_t)stream);
    GA_LAUNCH_CHECK();
    return GA_OK;
  }
  if (zring)
    hipLaunchKernelGGL(fields_alm_kernel<true>, grid, block, 0, (hipStream_t)stream, nl, k,
                       first_shell, nshell, nalm, w, reinterpret_cast<const double2 *>(zring),
                       nslot, seed, reinterpret_cast<float2 *>(out_f32), ldo,
                       reinterpret_cast<double2 *>(out_f64));
  else
    hipLaunchKernelGGL(fields_alm_kernel<false>, grid, block, 0, (hipStream_t)stream, nl, k,
                       first_shell, nshell, nalm, w, nullptr, 1, seed,
                       reinterpret_cast<float2 *>(out_f32), ldo,
                       reinterpret_cast<double2 *>(out_f64));
  GA_LAUNCH_CHECK();
  return GA_OK;
}

extern "C" int ga_fields_colmax_rows(int32_t nl, int64_t *rows) {
  clear_error();
  GA_REQUIRE(nl > 0 && rows, GA_ERR_INVALID, "invalid argument to ga_fields_colmax_rows");
  *rows = (int64_t(nl) * (nl + 1) / 2 + kFaBlock - 1) / kFaBlock;
  return GA_OK;
}

extern "C" int ga_fields_alm_colmax(int32_t nl, int32_t k, int32_t first_shell, int32_t nshell,
                                    const double *w, uint64_t seed, float *out_f32, int32_t ldo,
                                    float *colpart, void *stream) {
  clear_error();
  GA_REQUIRE(nl > 0 && k >= 0 && first_shell >= 0 && nshell > 0 && w && out_f32 && colpart,
             GA_ERR_INVALID, "invalid argument to ga_fields_alm_colmax");
  GA_REQUIRE(k + 1 <= kWinMax, GA_ERR_UNSUPPORTED,
             "column maxima need k + 1 <= %d neighbours (k = %d)", kWinMax, k);
  if (ldo == 0) ldo = nshell;
  GA_REQUIRE(ldo >= nshell, GA_ERR_INVALID, "ldo %d < nshell %d", ldo, nshell);
  const int64_t nalm = int64_t(nl) * (nl + 1) / 2;
  launch_window(nl, k, first_shell, nshell, nalm, w, seed, reinterpret_cast<float2 *>(out_f32),
                ldo, colpart, (hipStream_t)stream);
  GA_LAUNCH_CHECK();
  return GA_OK;
}

extern "C" int ga_multalm_f64(int32_t nl, const double *alm, const double *bl, double *out,
                              void *stream) {
  clear_error();
  GA_REQUIRE(nl >= 0 && (nl == 0 || (alm && bl && out)), GA_ERR_INVALID, "invalid argument");
  const int64_t n = int64_t(nl) * (nl + 1) / 2;
  if (n == 0) return GA_OK;
  hipLaunchKernelGGL(multalm_kernel, dim3(nblocks(n, 256)), dim3(256), 0, (hipStream_t)stream, n,
                     reinterpret_cast<const double2 *>(alm), bl, reinterpret_cast<double2 *>(out));
  GA_LAUNCH_CHECK();
  return GA_OK;
}

extern "C" int ga_glass_to_healpix_alm_f64(int32_t nl, const double *alm_glass, double *alm_hp,
                                           void *stream) {
  clear_error();
  GA_REQUIRE(nl > 0 && alm_glass && alm_hp && alm_glass != alm_hp, GA_ERR_INVALID,
             "invalid argument (in-place reorder is not supported)");
  const int64_t n = int64_t(nl) * (nl + 1) / 2;
  hipLaunchKernelGGL(glass_to_healpix_kernel, dim3(nblocks(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, nl, n, reinterpret_cast<const double2 *>(alm_glass),
                     reinterpret_cast<double2 *>(alm_hp));
  GA_LAUNCH_CHECK();
  return GA_OK;
}

extern "C" int ga_almxfl(int32_t lmax, int32_t nshell, int32_t real_bytes, const void *alm,
                         const double *fl, int32_t nfl, void *out, void *stream) {
  clear_error();
  GA_REQUIRE(lmax >= 0 && nshell > 0 && alm && out && (fl || nfl == 0) && nfl >= 0,
             GA_ERR_INVALID, "invalid argument to ga_almxfl");
  const int64_t n = nalm_of(lmax) * nshell;
  if (real_bytes == 8)
    hipLaunchKernelGGL(almxfl_kernel<double2>, dim3(nblocks(n, 256)), dim3(256), 0,
                       (hipStream_t)stream, lmax, nshell, n, (const double2 *)alm, fl, nfl,
                       (double2 *)out);
  else if (real_bytes == 4)
    hipLaunchKernelGGL(almxfl_kernel<float2>, dim3(nblocks(n, 256)), dim3(256), 0,
                       (hipStream_t)stream, lmax, nshell, n, (const float2 *)alm, fl, nfl,
                       (float2 *)out);
  else
    GA_REQUIRE(false, GA_ERR_INVALID, "real_bytes must be 4 or 8");
  GA_LAUNCH_CHECK();
  return GA_OK;
}

extern "C" int ga_transform_f64(int64_t n, const double *x, const ga_epilogue *epi, double *out,
                                void *stream) {
  clear_error();
  GA_REQUIRE(n >= 0 && epi && (n == 0 || (x && out)), GA_ERR_INVALID, "invalid argument");
  if (n == 0) return GA_OK;
  hipLaunchKernelGGL(transform_f64_kernel, dim3(nblocks(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, n, x, epi->kind, epi->lamda, epi->param, out);
  GA_LAUNCH_CHECK();
  return GA_OK;
}
