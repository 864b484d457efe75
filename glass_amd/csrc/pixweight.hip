// HEALPix pixel quadrature weights, computed on the GPU.
//
// healpy's map2alm(use_pixel_weights=True) (called at glass/lensing.py:306
// and :408 through glass/healpix.py:232-276) reads per-pixel weights from
// data files.  Here they are computed from the property that defines them:
// weights w_p for which the pixel sum is the integral over the sphere,
//     C(w)_lm = sum_p w_p Y*_lm(p) = sqrt(4 pi) delta_l0 delta_m0,
// for every l <= L (the plan's lmax), taken as the smallest correction of
// the uniform weights 4 pi / npix: w = w0 + S(lambda), where S is the
// synthesis (the adjoint of C under the inner product
// <y, z> = sum_l Re y_l0 z_l0 + 2 sum_{m > 0} Re y_lm conj(z_lm), for which
// S is the transpose of C), and lambda solves C S lambda = b - C(w0) by
// conjugate gradients.  C(f) = (npix / 4 pi) A(f) with A the plan's FP64
// analysis (one pass, uniform weights); S is the plan's FP64 synthesis.  With
// L <= 2 nside the pixels outnumber the conditions about three to one and the
// system is well conditioned.  The inner products are reduced over a fixed
// grid in a fixed order and summed on the host in order, so the weights are
// reproducible.
#include <vector>

#include "common.h"
#include "device.h"
#include "kernels.h"

namespace ga {
namespace {

constexpr int kDotBlocks = 1024;

// per-block partial sums of <y, z> (healpy order: the m = 0 entries are the
// first lmax + 1, weight 1; the others weight 2)
__global__ __launch_bounds__(256) void alm_dot_kernel(int64_t nalm, int lmax,
                                                      const double2 *__restrict__ y,
                                                      const double2 *__restrict__ z,
                                                      double *__restrict__ part) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < nalm;
       i += int64_t(gridDim.x) * 256) {
    const double f = i <= lmax ? 1.0 : 2.0;
    acc += f * (y[i].x * z[i].x + y[i].y * z[i].y);
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// y = c y, with Im y_l0 = 0 (the synthesis ignores it; C of a real map has none)
__global__ void alm_scale_kernel(int64_t nalm, int lmax, double c, double2 *__restrict__ y) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nalm) return;
  double2 v = y[i];
  v.x *= c;
  v.y = i <= lmax ? 0.0 : v.y * c;
  y[i] = v;
}

// r = b - r with b = sqrt(4 pi) at (0, 0)
__global__ void alm_rhs_kernel(int64_t nalm, int lmax, double2 *__restrict__ r) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nalm) return;
  double2 v = r[i];
  v.x = (i == 0 ? sqrt(4.0 * M_PI) : 0.0) - v.x;
  v.y = i <= lmax ? 0.0 : -v.y;
  r[i] = v;
}

// y += a x
__global__ void alm_axpy_kernel(int64_t nalm, double a, const double2 *__restrict__ x,
                                double2 *__restrict__ y) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nalm) return;
  double2 v = y[i];
  v.x = fma(a, x[i].x, v.x);
  v.y = fma(a, x[i].y, v.y);
  y[i] = v;
}

// p = r + b p
__global__ void alm_xpby_kernel(int64_t nalm, const double2 *__restrict__ r, double b,
                                double2 *__restrict__ p) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nalm) return;
  double2 v = p[i];
  v.x = fma(b, v.x, r[i].x);
  v.y = fma(b, v.y, r[i].y);
  p[i] = v;
}

__global__ void fill_kernel(int64_t n, double v, double *__restrict__ x) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// w = w0 + x
__global__ void add_const_kernel(int64_t n, double w0, const double *__restrict__ x,
                                 double *__restrict__ w) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) w[i] = w0 + x[i];
}

unsigned blocks_of(int64_t n) { return unsigned((n + 255) / 256); }

struct pw_layout {
  size_t lam, r, pv, q, map, part, sht, total;
};

pw_layout pw_workspace(const ga_plan *p) {
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  pw_layout l;
  const size_t av = up(size_t(p->nalm) * 16);
  l.lam = 0;
  l.r = l.lam + av;
  l.pv = l.r + av;
  l.q = l.pv + av;
  l.map = l.q + av;
  l.part = l.map + up(size_t(p->npix) * 8);
  l.sht = l.part + up(kDotBlocks * 8);
  size_t sht = 0;
  (void)ga_map2alm_workspace_size(p, 1, 8, &sht);
  l.total = l.sht + up(sht);
  return l;
}

}  // namespace
}  // namespace ga

using namespace ga;

extern "C" int ga_pixel_weights_workspace_size(const ga_plan *p, size_t *bytes) {
  clear_error();
  GA_REQUIRE(p && bytes, GA_ERR_INVALID, "NULL argument");
  *bytes = pw_workspace(p).total;
  return GA_OK;
}

extern "C" int ga_pixel_weights(const ga_plan *p, double tol, int32_t maxiter, double *weights,
                                int32_t *iterations, double *residual, void *ws,
                                size_t ws_bytes, void *stream) {
  clear_error();
  GA_REQUIRE(p && weights && ws && tol > 0.0 && maxiter >= 1, GA_ERR_INVALID,
             "invalid argument");
  GA_REQUIRE(p->lmax <= 2 * p->nside, GA_ERR_UNSUPPORTED,
             "pixel weights are solved for lmax <= 2 nside (plan lmax %d, nside %d)", p->lmax,
             p->nside);
  const pw_layout L = pw_workspace(p);
  GA_REQUIRE(ws_bytes >= L.total, GA_ERR_WORKSPACE, "workspace %zu < %zu bytes", ws_bytes,
             L.total);
  hipStream_t s = (hipStream_t)stream;
  char *base = static_cast<char *>(ws);
  double2 *lam = reinterpret_cast<double2 *>(base + L.lam);
  double2 *r = reinterpret_cast<double2 *>(base + L.r);
  double2 *pv = reinterpret_cast<double2 *>(base + L.pv);
  double2 *q = reinterpret_cast<double2 *>(base + L.q);
  double *map = reinterpret_cast<double *>(base + L.map);
  double *part = reinterpret_cast<double *>(base + L.part);
  void *sht = base + L.sht;
  size_t sht_bytes = 0;
  (void)ga_map2alm_workspace_size(p, 1, 8, &sht_bytes);
  const int64_t nalm = p->nalm, npix = p->npix;
  const int lmax = p->lmax;
  const double kappa = double(npix) / (4.0 * M_PI);  // C = kappa A
  std::vector<double> hp(kDotBlocks);
  int rc;

  auto dot = [&](const double2 *y, const double2 *z, double *out) -> int {
    hipLaunchKernelGGL(alm_dot_kernel, dim3(kDotBlocks), dim3(256), 0, s, nalm, lmax, y, z, part);
    GA_LAUNCH_CHECK();
    GA_CHECK_HIP(hipMemcpyAsync(hp.data(), part, kDotBlocks * sizeof(double),
                                hipMemcpyDeviceToHost, s));
    GA_CHECK_HIP(hipStreamSynchronize(s));
    double acc = 0.0;
    for (int i = 0; i < kDotBlocks; ++i) acc += hp[i];
    *out = acc;
    return GA_OK;
  };
  // q = C(f) for the map in `map`
  auto apply_c = [&](double2 *out) -> int {
    int e = ga_map2alm_f64(p, map, reinterpret_cast<double *>(out), 0, nullptr, sht, sht_bytes,
                           stream);
    if (e) return e;
    hipLaunchKernelGGL(alm_scale_kernel, dim3(blocks_of(nalm)), dim3(256), 0, s, nalm, lmax,
                       kappa, out);
    GA_LAUNCH_CHECK();
    return GA_OK;
  };
  auto apply_s = [&](const double2 *y) -> int {
    return ga_alm2map_f64(p, reinterpret_cast<const double *>(y), map, sht, sht_bytes, stream);
  };

  // r = b - C(w0), C(w0) = C(4 pi / npix) = A(1)
  hipLaunchKernelGGL(fill_kernel, dim3(blocks_of(npix)), dim3(256), 0, s, npix,
                     4.0 * M_PI / double(npix), map);
  GA_LAUNCH_CHECK();
  if ((rc = apply_c(r))) return rc;
  hipLaunchKernelGGL(alm_rhs_kernel, dim3(blocks_of(nalm)), dim3(256), 0, s, nalm, lmax, r);
  GA_LAUNCH_CHECK();
  hipLaunchKernelGGL(fill_kernel, dim3(blocks_of(2 * nalm)), dim3(256), 0, s, 2 * nalm, 0.0,
                     reinterpret_cast<double *>(lam));
  GA_LAUNCH_CHECK();
  GA_CHECK_HIP(hipMemcpyAsync(pv, r, size_t(nalm) * 16, hipMemcpyDeviceToDevice, s));
  double rr = 0.0, rr0 = 0.0;
  if ((rc = dot(r, r, &rr))) return rc;
  rr0 = rr;
  int it = 0;
  while (it < maxiter && rr > tol * tol * rr0 && rr > 0.0) {
    if ((rc = apply_s(pv))) return rc;
    if ((rc = apply_c(q))) return rc;
    double pq = 0.0;
    if ((rc = dot(pv, q, &pq))) return rc;
    GA_REQUIRE(pq > 0.0, GA_ERR_INVALID, "pixel weights: the normal operator is not positive");
    const double alpha = rr / pq;
    hipLaunchKernelGGL(alm_axpy_kernel, dim3(blocks_of(nalm)), dim3(256), 0, s, nalm, alpha, pv,
                       lam);
    hipLaunchKernelGGL(alm_axpy_kernel, dim3(blocks_of(nalm)), dim3(256), 0, s, nalm, -alpha, q,
                       r);
    GA_LAUNCH_CHECK();
    double rn = 0.0;
    if ((rc = dot(r, r, &rn))) return rc;
    hipLaunchKernelGGL(alm_xpby_kernel, dim3(blocks_of(nalm)), dim3(256), 0, s, nalm, r, rn / rr,
                       pv);
    GA_LAUNCH_CHECK();
    rr = rn;
    ++it;
  }
  // w = w0 + S(lambda)
  if ((rc = apply_s(lam))) return rc;
  hipLaunchKernelGGL(add_const_kernel, dim3(blocks_of(npix)), dim3(256), 0, s, npix,
                     4.0 * M_PI / double(npix), map, weights);
  GA_LAUNCH_CHECK();
  if (iterations) *iterations = it;
  if (residual) *residual = rr0 > 0.0 ? sqrt(rr / rr0) : 0.0;
  return GA_OK;
}
