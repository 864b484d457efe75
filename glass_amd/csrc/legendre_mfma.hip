// Batched Legendre stage of the scalar synthesis on FP32 MFMA (gfx950).
//
// For one m, the ring coefficients of S maps are a GEMM
//     F_m[ring, s] = sum_l lambda_lm(theta_ring) a_lm[s]
// with K = l.  Using lambda_lm(-x) = (-1)^(l+m) lambda_lm(x) the contraction is
// split by the parity of l - m into E (even) and O (odd), and one pass gives
// the north ring (E + O) and its south mirror (E - O).
//
// Work decomposition: a workgroup = 4 waves = 256 consecutive ring pairs of one
// m and one column group (32 NT columns of the [idx][S] complex64 alm viewed as
// 2S real columns).  Each lane owns one ring pair and runs the FP64 three-term
// recurrence (scaled seed, no underflow), four l per step ("quad" l0 .. l0+3).
// The four FP64 values are rounded to FP32 and fed straight into
// v_mfma_f32_32x32x2_f32 as the A operand: for parity p the K = 2 slice is
// (l0 + p, l0 + p + 2), and v_permlane32_swap moves the two values between the
// wave halves so that each 32-ring row tile gets [lambda(l0+p) | lambda(l0+p+2)]
// in lanes [0..31 | 32..63] -- the MFMA A layout -- without touching LDS.
// B is the alm slice (rows l0 + p, l0 + p + 2; 32 columns), read from HBM/L2
// with one coalesced 128-byte row per wave half.  Accumulation is FP32 in
// AGPRs.  While every lane of a wave is still below the Legendre turn-on (the
// seed sin^m theta far below FP32 range) the MFMAs are skipped.
#include "common.h"
#include "device.h"
#include "kernels.h"

namespace ga {

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float u2f(unsigned u) { return __uint_as_float(u); }
__device__ __forceinline__ unsigned f2u(float f) { return __float_as_uint(f); }

template <int NT>
__global__ __launch_bounds__(256, 2) void legendre_synth_mfma_kernel(
    int lmax, int npairs, int nrings, int ncols, int nrg, int ncg,
    const double *__restrict__ z, const double *__restrict__ sth, const double2 *__restrict__ ab,
    const double *__restrict__ cm, const float *__restrict__ alm, float *__restrict__ F) {
  // XCD-aware decode: block b runs on XCD b % 8; XCD x takes m = x, x + 8, ...
  // (largest l-range first), all (ring group, column group) tiles of one m
  // back to back so they share that m's alm slice in the XCD's L2.
  const int b = blockIdx.x;
  const int xcd = b & 7, kq = b >> 3, G = nrg * ncg;
  const int m = 8 * (kq / G) + xcd;
  if (m > lmax) return;
  const int rem = kq - (kq / G) * G;
  const int rg = rem / ncg, cg = rem - (rem / ncg) * ncg;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int rp0 = rg * 256 + wave * 64;
  if (rp0 >= npairs) return;  // wave-uniform; the kernel has no barriers
  const int rp = rp0 + lane;
  const int rpc = rp < npairs ? rp : npairs - 1;
  const double x = z[rpc];
  double p0, p1 = 0.0;  // lambda_{l-2}, lambda_{l-1}; the table entry at l = m is
  int k;                // (1, -1), so one step from (seed, 0) yields lambda_mm
  legendre_seed(cm[m], sth[rpc], m, p0, k);

  const int64_t base = hp_index(lmax, m, m);
  const double2 *__restrict__ co = ab + base;
  const int half = lane >> 5, col_in = lane & 31;
  const int colbase = cg * 32 * NT;
  const float *__restrict__ bptr = alm + base * ncols + colbase + col_in;
  bool colok[NT];
#pragma unroll
  for (int ct = 0; ct < NT; ++ct)
    colok[ct] = (colbase + ct * 32 + col_in < ncols) && !(m == 0 && (col_in & 1));

  f32x16 acc[2][2][NT];  // [parity][row tile][col tile]
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) acc[p][t][ct] = f32x16{};

  // One loop body for every quad so the accumulators stay in place.  `check`
  // (wave-uniform) is true while some lane still carries a scaled value below
  // the representable range; the MFMAs are skipped while no lane is in range.
  bool check = true;
  for (int l0 = m; l0 <= lmax; l0 += 4) {
    const bool tail = l0 + 3 > lmax;
    float f[4];
    if (check) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double2 c = co[l0 + j - m];
        const double pn = c.x * (x * p1 - c.y * p0);
        p0 = p1;
        p1 = pn;
        if (k < 0 && fabs(p1) > 0x1p300) { p1 *= 0x1p-600; p0 *= 0x1p-600; ++k; }
        f[j] = float(k == 0 ? p1 : 0.0);
      }
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double2 c = co[l0 + j - m];
        const double pn = c.x * (x * p1 - c.y * p0);
        p0 = p1;
        p1 = pn;
        f[j] = float(p1);
      }
    }
    if (tail) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (l0 + j > lmax) f[j] = 0.0f;
    }
    const bool on = !check || wave_any(k == 0);
    if (check) check = wave_any(k < 0);
    if (!on) continue;
    const auto sw0 = __builtin_amdgcn_permlane32_swap(f2u(f[0]), f2u(f[2]), false, false);
    const auto sw1 = __builtin_amdgcn_permlane32_swap(f2u(f[1]), f2u(f[3]), false, false);
    const float av[2][2] = {{u2f(sw0[0]), u2f(sw0[1])}, {u2f(sw1[0]), u2f(sw1[1])}};
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int l = l0 + p + 2 * half;
      const bool lok = l <= lmax;
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) {
        const float bv = (lok && colok[ct]) ? bptr[int64_t(l - m) * ncols + ct * 32] : 0.0f;
        acc[p][0][ct] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[p][0], bv, acc[p][0][ct], 0, 0, 0);
        acc[p][1][ct] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[p][1], bv, acc[p][1][ct], 0, 0, 0);
      }
    }
  }

  // epilogue: D layout of 32x32x2: lane L, acc element e -> row 8 (e/4) + 4 (L/32) + e%4,
  // column L % 32.  north = E + O, south = E - O.
  const int64_t mstride = int64_t(lmax + 1) * ncols;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) {
      const int col = colbase + ct * 32 + col_in;
      if (col >= ncols) continue;
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int row = 8 * (e >> 2) + 4 * half + (e & 3);
        const int rpo = rp0 + 32 * t + row;
        if (rpo >= npairs) continue;
        const float ev = acc[0][t][ct][e], ov = acc[1][t][ct][e];
        F[int64_t(rpo) * mstride + int64_t(m) * ncols + col] = ev + ov;
        if (rpo < npairs - 1) F[int64_t(nrings - 1 - rpo) * mstride + int64_t(m) * ncols + col] = ev - ov;
      }
    }
  }
}

}  // namespace

int launch_legendre_synth_mfma(const ga_plan *p, int S, const float *alm, float *F,
                               hipStream_t stream) {
  const int ncols = 2 * S;
  const int nt = ncols <= 32 ? 1 : 2;
  const int nrg = (p->npairs + 255) / 256;
  const int ncg = (ncols + 32 * nt - 1) / (32 * nt);
  const int nm8 = (p->lmax + 1 + 7) / 8;
  const int64_t nblk = int64_t(8) * nm8 * nrg * ncg;
  GA_REQUIRE(nblk < (int64_t(1) << 31), GA_ERR_UNSUPPORTED, "grid too large");
  if (nt == 1)
    hipLaunchKernelGGL(legendre_synth_mfma_kernel<1>, dim3(unsigned(nblk)), dim3(256), 0, stream,
                       p->lmax, p->npairs, p->nrings, ncols, nrg, ncg, p->d_z, p->d_sth, p->d_ab,
                       p->d_cm, alm, F);
  else
    hipLaunchKernelGGL(legendre_synth_mfma_kernel<2>, dim3(unsigned(nblk)), dim3(256), 0, stream,
                       p->lmax, p->npairs, p->nrings, ncols, nrg, ncg, p->d_z, p->d_sth, p->d_ab,
                       p->d_cm, alm, F);
  GA_LAUNCH_CHECK();
  return GA_OK;
}

}  // namespace ga
