// Positions tracing a density contrast (glass/points.py:443-540, with the
// per-pixel sampling of glass/healpix.py:398-432 randang).
//
// Per population: the expected number of points per pixel,
//     lambda_p = max(0, ((b(delta_p) [- mean]) + 1) * scale [* vis_p]),
// scale = ARCMIN2_SPHERE / npix * ngal, evaluated with the reference's
// roundings (glass/points.py:214-316; the bias model in the kernel for the
// linear and log-linear models); a Poisson count per pixel; then every point
// placed uniformly inside its pixel.  Counts and positions come from the
// counter-based Philox4x32-10 generator keyed by (seed, population, pixel or
// point), so they do not depend on launch shape or order:
//   * count: inversion of the Poisson CDF, lambda cut into parts of at most
//     32 (a sum of independent Poisson counts is Poisson), one uniform per
//     part from counter (pixel lo, pixel hi, population, 0x40000000 + part);
//   * position of point q: u, v from counter (q lo, q hi, population,
//     0x80000000) at face coordinates ((ix + u) / nside, (iy + v) / nside)
//     of the pixel, which are equal-area, so the points are uniform in area.
// The point offsets are a deterministic two-level scan: per block of 1024
// pixels a count, one workgroup scanning the block counts, and each block
// scanning its own pixels again when it writes its points.
#include "common.h"
#include "device.h"
#include "kernels.h"

namespace ga {
namespace {

constexpr int kPtThreads = 256;
constexpr int kPtPerThread = 4;
constexpr int kPtBlock = kPtThreads * kPtPerThread;  // pixels per block
constexpr int kPoissonPart = 32;                      // largest lambda per inversion
constexpr int kPoissonMaxK = 2048;                    // cap of one inversion

enum { kModelIdentity = 0, kModelLinear = 1, kModelLogLinear = 2 };

__constant__ int kPtJrll[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};
__constant__ int kPtJpll[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};

template <typename T>
__device__ __forceinline__ double biased(const T *__restrict__ delta, int64_t i, int model,
                                         double b) {
  const double d = double(delta[i]);
  if (model == kModelLinear) return b * d;
  if (model == kModelLogLinear) return expm1(log1p(d) * b);
  return d;
}

// per block: sum of the biased contrast (fixed order: thread partials, then a
// tree in LDS), for the monopole
template <typename T>
__global__ __launch_bounds__(kPtThreads) void points_sum_kernel(int64_t npix,
                                                                const T *__restrict__ delta,
                                                                int model, double b,
                                                                double *__restrict__ part) {
  __shared__ double red[kPtThreads];
  const int64_t base = int64_t(blockIdx.x) * kPtBlock;
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < kPtPerThread; ++j) {
    const int64_t i = base + int64_t(j) * kPtThreads + threadIdx.x;
    if (i < npix) acc += biased(delta, i, model, b);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kPtThreads / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// one workgroup: mean = sum of the block partials in order / npix
__global__ __launch_bounds__(1024) void points_mean_kernel(int64_t nblk, int64_t npix,
                                                           const double *__restrict__ part,
                                                           double *__restrict__ mean) {
  __shared__ double red[1024];
  const int64_t per = (nblk + 1023) / 1024;
  const int64_t lo = int64_t(threadIdx.x) * per, hi = min(nblk, lo + per);
  double acc = 0.0;
  for (int64_t i = lo; i < hi; ++i) acc += part[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *mean = red[0] / double(npix);
}

__device__ __forceinline__ int64_t poisson_count(double lam, uint64_t seed, int64_t pop,
                                                 int64_t pix) {
  int64_t n = 0;
  uint32_t part = 0;
  while (lam > 0.0) {
    const double l = fmin(lam, double(kPoissonPart));
    lam -= l;
    u32x4 c{{uint32_t(pix), uint32_t(uint64_t(pix) >> 32), uint32_t(pop), 0x40000000u + part}};
    const u32x4 r = philox4x32_10(c, uint32_t(seed), uint32_t(seed >> 32));
    ++part;
    const double u = uniform53(r.v[0], r.v[1]);
    double p = exp(-l), cdf = p;
    int k = 0;
    while (u > cdf && k < kPoissonMaxK) {
      ++k;
      p = p * (l / double(k));
      cdf = cdf + p;
    }
    n += k;
  }
  return n;
}

// expected counts (lam, optional) and Poisson counts (counts, optional), and
// the count of each block of 1024 pixels (blk, with counts)
template <typename T>
__global__ __launch_bounds__(kPtThreads) void points_count_kernel(
    int64_t npix, const T *__restrict__ delta, int model, double b, const double *__restrict__ mean,
    double scale, const double *__restrict__ vis, uint64_t seed, int64_t pop,
    double *__restrict__ lam_out, int64_t *__restrict__ counts, int64_t *__restrict__ blk) {
  // every product and sum rounded on its own, as numpy evaluates them
#pragma clang fp contract(off)
  __shared__ int64_t red[kPtThreads];
  const int64_t base = int64_t(blockIdx.x) * kPtBlock;
  const double mu = mean ? *mean : 0.0;
  int64_t tot = 0;
#pragma unroll
  for (int j = 0; j < kPtPerThread; ++j) {
    const int64_t i = base + int64_t(j) * kPtThreads + threadIdx.x;
    if (i >= npix) continue;
    double n = biased(delta, i, model, b);
    if (mean) n = n - mu;
    n = n + 1.0;
    n = n * scale;
    if (vis) n = n * vis[i];
    n = fmax(n, 0.0);
    if (lam_out) lam_out[i] = n;
    if (counts) {
      const int64_t c = poisson_count(n, seed, pop, i);
      counts[i] = c;
      tot += c;
    }
  }
  if (!counts) return;
  red[threadIdx.x] = tot;
  __syncthreads();
  for (int w = kPtThreads / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) blk[blockIdx.x] = red[0];
}

// block counts of given per-pixel counts (the host-generator path)
__global__ __launch_bounds__(kPtThreads) void points_block_count_kernel(
    int64_t npix, const int64_t *__restrict__ counts, int64_t *__restrict__ blk) {
  __shared__ int64_t red[kPtThreads];
  const int64_t base = int64_t(blockIdx.x) * kPtBlock;
  int64_t tot = 0;
#pragma unroll
  for (int j = 0; j < kPtPerThread; ++j) {
    const int64_t i = base + int64_t(j) * kPtThreads + threadIdx.x;
    if (i < npix) tot += counts[i];
  }
  red[threadIdx.x] = tot;
  __syncthreads();
  for (int w = kPtThreads / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) blk[blockIdx.x] = red[0];
}

// one workgroup: exclusive scan of the block counts in place, total at the end
__global__ __launch_bounds__(1024) void points_scan_kernel(int64_t nblk, int64_t *__restrict__ blk,
                                                           int64_t *__restrict__ total) {
  __shared__ int64_t sc[1024];
  const int64_t per = (nblk + 1023) / 1024;
  const int64_t lo = int64_t(threadIdx.x) * per, hi = min(nblk, lo + per);
  int64_t acc = 0;
  for (int64_t i = lo; i < hi; ++i) acc += blk[i];
  sc[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan
    const int64_t v = int(threadIdx.x) >= off ? sc[threadIdx.x - off] : 0;
    __syncthreads();
    sc[threadIdx.x] += v;
    __syncthreads();
  }
  int64_t run = sc[threadIdx.x] - acc;  // exclusive prefix of this thread's range
  for (int64_t i = lo; i < hi; ++i) {
    const int64_t c = blk[i];
    blk[i] = run;
    run += c;
  }
  if (threadIdx.x == 1023) *total = sc[1023];
}

// RING pixel -> (face, ix, iy) (Gorski et al. 2005 ring / face geometry)
__device__ void ring_to_xyf(int64_t nside, int64_t pix, int &face, int64_t &ix, int64_t &iy) {
  const int64_t ncap = 2 * nside * (nside - 1), npix = 12 * nside * nside, nl2 = 2 * nside;
  int64_t iring, iphi, kshift, nr;
  if (pix < ncap) {  // north polar cap
    iring = (1 + int64_t(sqrt(double(1 + 2 * pix)))) >> 1;
    while (2 * iring * (iring - 1) > pix) --iring;
    while (2 * (iring + 1) * iring <= pix) ++iring;
    iphi = (pix + 1) - 2 * iring * (iring - 1);
    kshift = 0;
    nr = iring;
    face = int((iphi - 1) / nr);
  } else if (pix < npix - ncap) {  // equatorial belt
    const int64_t ip = pix - ncap;
    const int64_t tmp = ip / (4 * nside);
    iring = tmp + nside;
    iphi = ip - tmp * 4 * nside + 1;
    kshift = (iring + nside) & 1;
    nr = nside;
    const int64_t ire = iring - nside + 1, irm = nl2 + 2 - ire;
    const int64_t ifm = (iphi - ire / 2 + nside - 1) / nside;
    const int64_t ifp = (iphi - irm / 2 + nside - 1) / nside;
    face = int(ifp == ifm ? (ifp | 4) : (ifp < ifm ? ifp : ifm + 8));
  } else {  // south polar cap
    const int64_t ip = npix - pix;
    int64_t ir = (1 + int64_t(sqrt(double(2 * ip - 1)))) >> 1;
    while (2 * ir * (ir - 1) >= ip) --ir;
    while (2 * (ir + 1) * ir < ip) ++ir;
    iphi = 4 * ir + 1 - (ip - 2 * ir * (ir - 1));
    kshift = 0;
    nr = ir;
    iring = 2 * nl2 - ir;
    face = int(8 + (iphi - 1) / nr);
  }
  const int64_t irt = iring - int64_t(kPtJrll[face]) * nside + 1;
  int64_t ipt = 2 * iphi - int64_t(kPtJpll[face]) * nr - kshift - 1;
  if (ipt >= nl2) ipt -= 8 * nside;
  ix = (ipt - irt) >> 1;
  iy = (-ipt - irt) >> 1;
}

// (lon, lat) in degrees of face coordinates (x, y) (equal-area HEALPix
// projection; lon in [0, 360))
__device__ void face_lonlat(int f, double x, double y, double &lon, double &lat) {
  const double jr = double(kPtJrll[f]) - x - y;
  double z, st, nr;
  if (jr < 1.0) {
    const double t = jr * jr / 3.0;
    z = 1.0 - t;
    st = sqrt(t * (2.0 - t));
    nr = jr;
  } else if (jr > 3.0) {
    nr = 4.0 - jr;
    const double t = nr * nr / 3.0;
    z = t - 1.0;
    st = sqrt(t * (2.0 - t));
  } else {
    z = (2.0 - jr) * (2.0 / 3.0);
    st = sqrt((1.0 - z) * (1.0 + z));
    nr = 1.0;
  }
  double q = nr > 0.0 ? (double(kPtJpll[f]) * nr + x - y) / nr : 0.0;  // phi in units of pi / 4
  q = q - 8.0 * floor(q / 8.0);
  lon = 45.0 * q;
  lat = atan2(z, st) * (180.0 / M_PI);
}

// points of each pixel at its offset (block offsets from points_scan_kernel
// plus the pixel's exclusive prefix inside the block)
__global__ __launch_bounds__(kPtThreads) void points_place_kernel(
    int64_t nside, int64_t npix, const int64_t *__restrict__ counts,
    const int64_t *__restrict__ blk, uint64_t seed, int64_t pop, double *__restrict__ lon,
    double *__restrict__ lat) {
  __shared__ int64_t sc[kPtThreads];
  const int64_t base = int64_t(blockIdx.x) * kPtBlock;
  // thread t owns pixels base + 4 t .. + 3 (contiguous, so that its prefix
  // is the scan of the thread sums)
  int64_t c[kPtPerThread];
  int64_t mine = 0;
#pragma unroll
  for (int j = 0; j < kPtPerThread; ++j) {
    const int64_t i = base + int64_t(threadIdx.x) * kPtPerThread + j;
    c[j] = i < npix ? counts[i] : 0;
    mine += c[j];
  }
  sc[threadIdx.x] = mine;
  __syncthreads();
  for (int off = 1; off < kPtThreads; off <<= 1) {
    const int64_t v = int(threadIdx.x) >= off ? sc[threadIdx.x - off] : 0;
    __syncthreads();
    sc[threadIdx.x] += v;
    __syncthreads();
  }
  int64_t q = blk[blockIdx.x] + sc[threadIdx.x] - mine;
  const double inv = 1.0 / double(nside);
#pragma unroll
  for (int j = 0; j < kPtPerThread; ++j) {
    const int64_t i = base + int64_t(threadIdx.x) * kPtPerThread + j;
    if (c[j] == 0) continue;
    int face;
    int64_t ix, iy;
    ring_to_xyf(nside, i, face, ix, iy);
    for (int64_t e = 0; e < c[j]; ++e, ++q) {
      u32x4 ctr{{uint32_t(q), uint32_t(uint64_t(q) >> 32), uint32_t(pop), 0x80000000u}};
      const u32x4 r = philox4x32_10(ctr, uint32_t(seed), uint32_t(seed >> 32));
      const double u = uniform53(r.v[0], r.v[1]), v = uniform53(r.v[2], r.v[3]);
      double lo, la;
      face_lonlat(face, (double(ix) + u) * inv, (double(iy) + v) * inv, lo, la);
      lon[q] = lo;
      lat[q] = la;
    }
  }
}

int64_t nblocks(int64_t npix) { return (npix + kPtBlock - 1) / kPtBlock; }

}  // namespace
}  // namespace ga

using namespace ga;

extern "C" int ga_points_workspace_size(int64_t npix, size_t *bytes) {
  clear_error();
  GA_REQUIRE(bytes && npix >= 1, GA_ERR_INVALID, "invalid argument");
  // block partials (double) and block counts / offsets (int64), the mean and the total
  *bytes = size_t(nblocks(npix)) * 16 + 256;
  return GA_OK;
}

extern "C" int ga_points_counts(int64_t npix, int32_t real_bytes, const void *delta, int32_t model,
                                double bias, int32_t remove_monopole, double scale,
                                const double *vis, uint64_t seed, int64_t population,
                                double *lam, int64_t *counts, int64_t *total, void *ws,
                                size_t ws_bytes, void *stream) {
  clear_error();
  GA_REQUIRE(npix >= 1 && delta && ws && (lam || counts), GA_ERR_INVALID, "invalid argument");
  GA_REQUIRE(real_bytes == 4 || real_bytes == 8, GA_ERR_INVALID, "real_bytes must be 4 or 8");
  GA_REQUIRE(model >= kModelIdentity && model <= kModelLogLinear, GA_ERR_INVALID,
             "unknown bias model %d", model);
  GA_REQUIRE(!counts || total, GA_ERR_INVALID, "counts need a total");
  size_t need = 0;
  (void)ga_points_workspace_size(npix, &need);
  GA_REQUIRE(ws_bytes >= need, GA_ERR_WORKSPACE, "workspace %zu < %zu bytes", ws_bytes, need);
  hipStream_t s = (hipStream_t)stream;
  const int64_t nb = nblocks(npix);
  GA_REQUIRE(nb < (int64_t(1) << 31), GA_ERR_UNSUPPORTED, "map too large");
  double *part = static_cast<double *>(ws);
  int64_t *blk = reinterpret_cast<int64_t *>(part + nb);
  double *mean = reinterpret_cast<double *>(blk + nb);
  if (remove_monopole) {
    if (real_bytes == 8)
      hipLaunchKernelGGL(points_sum_kernel<double>, dim3(unsigned(nb)), dim3(kPtThreads), 0, s,
                         npix, (const double *)delta, model, bias, part);
    else
      hipLaunchKernelGGL(points_sum_kernel<float>, dim3(unsigned(nb)), dim3(kPtThreads), 0, s,
                         npix, (const float *)delta, model, bias, part);
    GA_LAUNCH_CHECK();
    hipLaunchKernelGGL(points_mean_kernel, dim3(1), dim3(1024), 0, s, nb, npix, part, mean);
    GA_LAUNCH_CHECK();
  }
  const double *mp = remove_monopole ? mean : nullptr;
  if (real_bytes == 8)
    hipLaunchKernelGGL(points_count_kernel<double>, dim3(unsigned(nb)), dim3(kPtThreads), 0, s,
                       npix, (const double *)delta, model, bias, mp, scale, vis, seed, population,
                       lam, counts, blk);
  else
    hipLaunchKernelGGL(points_count_kernel<float>, dim3(unsigned(nb)), dim3(kPtThreads), 0, s,
                       npix, (const float *)delta, model, bias, mp, scale, vis, seed, population,
                       lam, counts, blk);
  GA_LAUNCH_CHECK();
  if (counts) {
    hipLaunchKernelGGL(points_scan_kernel, dim3(1), dim3(1024), 0, s, nb, blk, total);
    GA_LAUNCH_CHECK();
  }
  return GA_OK;
}

extern "C" int ga_points_scan(int64_t npix, const int64_t *counts, int64_t *total, void *ws,
                              size_t ws_bytes, void *stream) {
  clear_error();
  GA_REQUIRE(npix >= 1 && counts && total && ws, GA_ERR_INVALID, "invalid argument");
  size_t need = 0;
  (void)ga_points_workspace_size(npix, &need);
  GA_REQUIRE(ws_bytes >= need, GA_ERR_WORKSPACE, "workspace %zu < %zu bytes", ws_bytes, need);
  hipStream_t s = (hipStream_t)stream;
  const int64_t nb = nblocks(npix);
  GA_REQUIRE(nb < (int64_t(1) << 31), GA_ERR_UNSUPPORTED, "map too large");
  int64_t *blk = reinterpret_cast<int64_t *>(static_cast<double *>(ws) + nb);
  hipLaunchKernelGGL(points_block_count_kernel, dim3(unsigned(nb)), dim3(kPtThreads), 0, s, npix,
                     counts, blk);
  GA_LAUNCH_CHECK();
  hipLaunchKernelGGL(points_scan_kernel, dim3(1), dim3(1024), 0, s, nb, blk, total);
  GA_LAUNCH_CHECK();
  return GA_OK;
}

extern "C" int ga_points_positions(int32_t nside, const int64_t *counts, uint64_t seed,
                                   int64_t population, double *lon, double *lat, void *ws,
                                   size_t ws_bytes, void *stream) {
  clear_error();
  GA_REQUIRE(nside >= 1 && counts && lon && lat && ws, GA_ERR_INVALID, "invalid argument");
  const int64_t npix = 12 * int64_t(nside) * nside;
  size_t need = 0;
  (void)ga_points_workspace_size(npix, &need);
  GA_REQUIRE(ws_bytes >= need, GA_ERR_WORKSPACE, "workspace %zu < %zu bytes", ws_bytes, need);
  const int64_t nb = nblocks(npix);
  const int64_t *blk = reinterpret_cast<const int64_t *>(static_cast<const double *>(ws) + nb);
  hipLaunchKernelGGL(points_place_kernel, dim3(unsigned(nb)), dim3(kPtThreads), 0,
                     (hipStream_t)stream, int64_t(nside), npix, counts, blk, seed, population, lon,
                     lat);
  GA_LAUNCH_CHECK();
  return GA_OK;
}
