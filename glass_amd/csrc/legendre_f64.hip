// Plain FP64 Legendre stage of the scalar synthesis (one map, no MFMA).
//
// One lane per (m, ring pair).  The lane runs the normalised associated-
// Legendre recurrence in l for its ring's cos(theta) in FP64 (seed
// lambda_mm = c_m sin^m theta carried with a 2^600-step scale index so it never
// underflows), and accumulates the l-sums of a_lm lambda_lm split by the parity
// of l - m, which gives the north ring and its south mirror at once:
//     F_m(theta)      = E + O,   F_m(pi - theta) = E - O,
// because lambda_lm(-x) = (-1)^(l+m) lambda_lm(x).
// This is the honest FP64 baseline the batched FP32-MFMA path is compared to.
#include "common.h"
#include "device.h"
#include "kernels.h"

namespace ga {

namespace {

__global__ __launch_bounds__(256) void legendre_synth_f64_kernel(
    int lmax, int npairs, int nrings, const double *__restrict__ z,
    const double *__restrict__ sth, const double2 *__restrict__ ab,
    const double *__restrict__ cm, const int32_t *__restrict__ ring_pnum,
    const int32_t *__restrict__ ring_pden, const double2 *__restrict__ alm,
    double2 *__restrict__ F) {
  const int m = blockIdx.x;
  const int rp = blockIdx.y * blockDim.x + threadIdx.x;
  const int rpc = rp < npairs ? rp : npairs - 1;
  const double x = z[rpc];
  double p0 = 0.0, p1;
  int k;
  legendre_seed(cm[m], sth[rpc], m, p1, k);

  const int64_t base = hp_index(lmax, m, m);
  const double2 *__restrict__ a = alm + base;  // a[l - m]
  const double2 *__restrict__ co = ab + base;  // co[l - m]
  double er = 0.0, ei = 0.0, orr = 0.0, oi = 0.0;

  int l = m;
  // phase 1: some lane of the wave is still below the representable range
  for (;;) {
    if (l > m) {
      const double2 c = co[l - m];
      const double pn = c.x * (x * p1 - c.y * p0);
      p0 = p1;
      p1 = pn;
    }
    if (k < 0 && fabs(p1) > 0x1p300) { p1 *= 0x1p-600; p0 *= 0x1p-600; ++k; }
    const double ve = (k == 0) ? p1 : 0.0;
    double vo = 0.0;
    if (l + 1 <= lmax) {
      const double2 c = co[l + 1 - m];
      const double pn = c.x * (x * p1 - c.y * p0);
      p0 = p1;
      p1 = pn;
      if (k < 0 && fabs(p1) > 0x1p300) { p1 *= 0x1p-600; p0 *= 0x1p-600; ++k; }
      vo = (k == 0) ? p1 : 0.0;
    }
    if (wave_any(k == 0)) {
      const double2 ae = a[l - m];
      er = fma(ve, ae.x, er);
      ei = fma(ve, ae.y, ei);
      if (l + 1 <= lmax) {
        const double2 ao = a[l + 1 - m];
        orr = fma(vo, ao.x, orr);
        oi = fma(vo, ao.y, oi);
      }
    }
    l += 2;
    if (l > lmax || !wave_any(k < 0)) break;
  }
  // phase 2: every lane in range, (l - m) even at the top of the loop
  for (; l + 1 <= lmax; l += 2) {
    const double2 c0 = co[l - m], c1 = co[l + 1 - m];
    const double2 ae = a[l - m], ao = a[l + 1 - m];
    double pn = c0.x * (x * p1 - c0.y * p0);
    p0 = p1;
    p1 = pn;
    er = fma(p1, ae.x, er);
    ei = fma(p1, ae.y, ei);
    pn = c1.x * (x * p1 - c1.y * p0);
    p0 = p1;
    p1 = pn;
    orr = fma(p1, ao.x, orr);
    oi = fma(p1, ao.y, oi);
  }
  if (l <= lmax) {
    const double2 c0 = co[l - m];
    const double2 ae = a[l - m];
    const double pn = c0.x * (x * p1 - c0.y * p0);
    p1 = pn;
    er = fma(p1, ae.x, er);
    ei = fma(p1, ae.y, ei);
  }
  if (m == 0) { ei = 0.0; oi = 0.0; }  // Im a_l0 is ignored (real map)
  if (rp >= npairs) return;
  // ring phase e^{i m phi0}, phi0 = 2 pi pnum / pden (the same on the south mirror)
  const int pnum = ring_pnum[rp], pden = ring_pden[rp];
  double ps = 0.0, pc = 1.0;
  if (pnum) sincospi(2.0 * double((int64_t(m) * pnum) % pden) / double(pden), &ps, &pc);
  const int64_t stride = lmax + 1;
  const double nr = er + orr, ni = ei + oi, sr = er - orr, si = ei - oi;
  F[int64_t(rp) * stride + m] = double2{nr * pc - ni * ps, nr * ps + ni * pc};
  if (rp < npairs - 1)
    F[int64_t(nrings - 1 - rp) * stride + m] = double2{sr * pc - si * ps, sr * ps + si * pc};
}

}  // namespace

int launch_legendre_synth_f64(const ga_plan *p, const double *alm, double *F,
                              hipStream_t stream) {
  dim3 grid(unsigned(p->lmax + 1), unsigned((p->npairs + 255) / 256));
  hipLaunchKernelGGL(legendre_synth_f64_kernel, grid, dim3(256), 0, stream, p->lmax, p->npairs,
                     p->nrings, p->d_z, p->d_sth, p->d_ab, p->d_cm, p->d_ring_pnum,
                     p->d_ring_pden, reinterpret_cast<const double2 *>(alm),
                     reinterpret_cast<double2 *>(F));
  GA_LAUNCH_CHECK();
  return GA_OK;
}

}  // namespace ga
