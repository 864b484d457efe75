// Angular power spectrum <-> angular correlation function on Gauss-Legendre
// nodes, the transforms of the Gaussian-spectra solver (glass/grf/_solver.py
// :27-148 and glass/grf/_core.py:141-184, where the transformcl package does
// them).
//
//   corr(x_k) = sum_l (2l+1) / (4 pi) cl_l P_l(x_k)                (cl2corr)
//   cl_l      = 2 pi sum_k w_k corr(x_k) P_l(x_k)                   (corr2cl)
//
// With n Gauss-Legendre nodes (x_k, w_k) and n coefficients l < n the pair is
// exact (products P_l P_l' of degree < 2n - 1 are integrated exactly), so
// corr2cl(cl2corr(cl)) = cl to rounding.  P_l by its three-term recurrence in
// FP64, one thread per node; corr2cl reduces each l over the 64 nodes of a
// wave (xor butterfly) and then over waves in a fixed order, so the result
// does not depend on scheduling.  The nodes come from Newton's method on P_n
// from the asymptotic first guess cos(pi (i + 3/4) / (n + 1/2)), one thread
// per node pair (x, -x).
#include "common.h"
#include "device.h"
#include "kernels.h"

namespace ga {
namespace {

__global__ __launch_bounds__(256) void gl_nodes_kernel(int n, double *__restrict__ x,
                                                       double *__restrict__ w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (n + 1) / 2) return;
  double z = cos(M_PI * (double(i) + 0.75) / (double(n) + 0.5));
  double dp = 1.0;
  for (int it = 0; it < 100; ++it) {
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= n; ++k) {
      const double p2 = (double(2 * k - 1) * z * p1 - double(k - 1) * p0) / double(k);
      p0 = p1;
      p1 = p2;
    }
    dp = double(n) * (z * p1 - p0) / (z * z - 1.0);
    const double dz = p1 / dp;
    z -= dz;
    if (fabs(dz) < 1e-13) break;  // quadratic convergence: the step left ~dz^2
  }
  {  // derivative at the converged node
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= n; ++k) {
      const double p2 = (double(2 * k - 1) * z * p1 - double(k - 1) * p0) / double(k);
      p0 = p1;
      p1 = p2;
    }
    dp = double(n) * (z * p1 - p0) / (z * z - 1.0);
  }
  const double wi = 2.0 / ((1.0 - z * z) * dp * dp);
  // ascending order: x[n - 1 - i] = +z (z > 0 for i < n / 2), x[i] = -z
  x[n - 1 - i] = z;
  w[n - 1 - i] = wi;
  x[i] = -z;
  w[i] = wi;
  if ((n & 1) && i == n / 2) x[i] = 0.0;
}

__global__ __launch_bounds__(256) void cl2corr_kernel(int nl, const double *__restrict__ cl,
                                                      int n, const double *__restrict__ x,
                                                      double *__restrict__ corr) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double xk = x[k];
  double p0 = 1.0, p1 = xk;
  double acc = cl[0] * (1.0 / (4.0 * M_PI));
  if (nl > 1) acc += cl[1] * (3.0 / (4.0 * M_PI)) * xk;
  for (int l = 2; l < nl; ++l) {
    const double p2 = (double(2 * l - 1) * xk * p1 - double(l - 1) * p0) / double(l);
    p0 = p1;
    p1 = p2;
    acc += cl[l] * (double(2 * l + 1) / (4.0 * M_PI)) * p2;
  }
  corr[k] = acc;
}

// part[wave][l]: 2 pi sum over the wave's nodes of w_k corr_k P_l(x_k)
__global__ __launch_bounds__(256) void corr2cl_kernel(int n, const double *__restrict__ x,
                                                      const double *__restrict__ w,
                                                      const double *__restrict__ corr, int nl,
                                                      int nwaves, double *__restrict__ part) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int wave = k >> 6, lane = k & 63;
  if (wave >= nwaves) return;  // whole waves past the last row
  const bool ok = k < n;
  const double xk = ok ? x[k] : 0.0;
  const double a = ok ? 2.0 * M_PI * w[k] * corr[k] : 0.0;
  double p0 = 1.0, p1 = xk;
  double *__restrict__ out = part + int64_t(wave) * nl;
  for (int l = 0; l < nl; ++l) {
    double P;
    if (l == 0) {
      P = 1.0;
    } else if (l == 1) {
      P = xk;
    } else {
      P = (double(2 * l - 1) * xk * p1 - double(l - 1) * p0) / double(l);
      p0 = p1;
      p1 = P;
    }
    double v = a * P;
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) out[l] = v;
  }
}

__global__ void corr2cl_sum_kernel(int nl, int nwaves, const double *__restrict__ part,
                                   double *__restrict__ cl) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nl) return;
  double acc = 0.0;
  for (int wv = 0; wv < nwaves; ++wv) acc += part[int64_t(wv) * nl + l];
  cl[l] = acc;
}

}  // namespace
}  // namespace ga

using namespace ga;

extern "C" int ga_legendre_nodes(int32_t n, double *x, double *w, void *stream) {
  clear_error();
  GA_REQUIRE(n >= 1 && x && w, GA_ERR_INVALID, "invalid argument");
  const int half = (n + 1) / 2;
  hipLaunchKernelGGL(gl_nodes_kernel, dim3((half + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     n, x, w);
  GA_LAUNCH_CHECK();
  return GA_OK;
}

extern "C" int ga_cl2corr(int32_t nl, const double *cl, int32_t n, const double *x, double *corr,
                          void *stream) {
  clear_error();
  GA_REQUIRE(nl >= 1 && n >= 1 && cl && x && corr, GA_ERR_INVALID, "invalid argument");
  hipLaunchKernelGGL(cl2corr_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, nl,
                     cl, n, x, corr);
  GA_LAUNCH_CHECK();
  return GA_OK;
}

extern "C" int ga_corr2cl_workspace_size(int32_t n, int32_t nl, size_t *bytes) {
  clear_error();
  GA_REQUIRE(bytes && n >= 1 && nl >= 1, GA_ERR_INVALID, "invalid argument");
  *bytes = size_t((n + 63) / 64) * size_t(nl) * sizeof(double);
  return GA_OK;
}

extern "C" int ga_corr2cl(int32_t n, const double *x, const double *w, const double *corr,
                          int32_t nl, double *cl, void *ws, size_t ws_bytes, void *stream) {
  clear_error();
  GA_REQUIRE(n >= 1 && nl >= 1 && x && w && corr && cl && ws, GA_ERR_INVALID, "invalid argument");
  const int nwaves = (n + 63) / 64;
  const size_t need = size_t(nwaves) * size_t(nl) * sizeof(double);
  GA_REQUIRE(ws_bytes >= need, GA_ERR_WORKSPACE, "workspace %zu < %zu bytes", ws_bytes, need);
  hipStream_t s = (hipStream_t)stream;
  double *part = static_cast<double *>(ws);
  hipLaunchKernelGGL(corr2cl_kernel, dim3((nwaves * 64 + 255) / 256), dim3(256), 0, s, n, x, w,
                     corr, nl, nwaves, part);
  GA_LAUNCH_CHECK();
  hipLaunchKernelGGL(corr2cl_sum_kernel, dim3((nl + 255) / 256), dim3(256), 0, s, nl, nwaves, part,
                     cl);
  GA_LAUNCH_CHECK();
  return GA_OK;
}
