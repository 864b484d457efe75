// Multi-plane convergence update on device-resident maps.
//
// glass/lensing.py:511-586 (MultiPlaneConvergence.add_plane) forms the next
// convergence plane from the previous two and the mass plane between them:
//     kappa3 = (1 - t) kappa1 + t kappa2 + f delta2
// as three in-place array operations in float64 (numpy promotes a float32 map
// times a float64 scalar to float64 and casts back on the in-place store).
// The kernel restates the same three steps with the same roundings: each
// product and sum rounded separately (no fused multiply-add), and for float32
// maps the value rounded back to float32 after each step, so the result is
// bit-identical to the reference's arithmetic.  One pass over the maps:
// reads kappa1, kappa2, delta2 and writes kappa3 (which may alias kappa1).
#include "common.h"
#include "device.h"
#include "kernels.h"

namespace ga {
namespace {

template <typename T>
__global__ __launch_bounds__(256) void multi_plane_kernel(int64_t n, const T *__restrict__ k1,
                                                          const T *__restrict__ k2,
                                                          const T *__restrict__ d2, double t,
                                                          double f, T *out) {
  // every product and sum rounded on its own, as numpy evaluates them (the
  // pragma covers the operators written here; it does not reach into inlined
  // library helpers, so none are used)
#pragma clang fp contract(off)
  const double omt = 1.0 - t;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    // kappa3 *= 1 - t
    T r = T(double(k1[i]) * omt);
    // kappa3 += t * kappa2
    const double tk = t * double(k2[i]);
    r = T(double(r) + tk);
    // kappa3 += f * delta2 (delta2 = 0 before the first mass plane)
    const double fd = f * (d2 ? double(d2[i]) : 0.0);
    r = T(double(r) + fd);
    out[i] = r;
  }
}

}  // namespace
}  // namespace ga

using namespace ga;

extern "C" int ga_multi_plane_update(int64_t n, int32_t real_bytes, const void *kappa1,
                                     const void *kappa2, const void *delta2, double t, double f,
                                     void *kappa3, void *stream) {
  clear_error();
  GA_REQUIRE(n >= 0 && kappa1 && kappa2 && kappa3, GA_ERR_INVALID, "invalid argument");
  GA_REQUIRE(real_bytes == 4 || real_bytes == 8, GA_ERR_INVALID, "real_bytes must be 4 or 8");
  if (n == 0) return GA_OK;
  hipStream_t s = (hipStream_t)stream;
  const unsigned grid = unsigned(std::min<int64_t>((n + 255) / 256, 8192));
  if (real_bytes == 8)
    hipLaunchKernelGGL(multi_plane_kernel<double>, dim3(grid), dim3(256), 0, s, n,
                       (const double *)kappa1, (const double *)kappa2, (const double *)delta2, t,
                       f, (double *)kappa3);
  else
    hipLaunchKernelGGL(multi_plane_kernel<float>, dim3(grid), dim3(256), 0, s, n,
                       (const float *)kappa1, (const float *)kappa2, (const float *)delta2, t, f,
                       (float *)kappa3);
  GA_LAUNCH_CHECK();
  return GA_OK;
}
