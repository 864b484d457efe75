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

// CHUNK > 0: the MFMA accumulators are flushed into FP32 master sums every
// CHUNK quads, so each rounding acts on a partial sum of at most 2 CHUNK terms
// per parity instead of the full l-range (error ~ sqrt(K) -> ~K^(1/4)).
template <int NT, int CHUNK, int MINW>
__global__ __launch_bounds__(256, MINW) void legendre_synth_mfma_kernel(
    int lmax, int npairs, int nrings, int ncols, int nrg, int ncg,
    const double *__restrict__ z, const double *__restrict__ sth, const double2 *__restrict__ ab,
    const double *__restrict__ cm, const int32_t *__restrict__ ring_pnum,
    const int32_t *__restrict__ ring_pden, const float *__restrict__ alm, float *__restrict__ F) {
  // XCD-aware decode: block b runs on XCD b % 8; XCD x takes the runs of four
  // consecutive m starting at 4 x, 4 x + 32, ... (largest l-range first; the
  // four m of a 128-byte line of the ring buffer go through one XCD's L2), all
  // (ring group, column group) tiles of one m back to back so they share that
  // m's alm slice in the XCD's L2.
  const int b = blockIdx.x;
  const int xcd = b & 7, kq = b >> 3, G = nrg * ncg;
  const int mq = kq / G;
  const int m = 32 * (mq >> 2) + 4 * xcd + (mq & 3);
  if (m > lmax) return;
  const int rem = kq - mq * G;
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
  // per-lane constant offsets of the B elements inside a quad's rows (rows l0 + p + 2 half):
  // columns past ncols read row-start data instead (their outputs are never
  // stored), so every load is in bounds and needs no mask
  int boff[2][NT];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) {
      const int col = colbase + ct * 32 + col_in;
      boff[p][ct] = (col < ncols) ? (p + 2 * half) * ncols + col : (p + 2 * half) * ncols;
    }

  f32x16 acc[2][2][NT];  // [parity][row tile][col tile]
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) acc[p][t][ct] = f32x16{};
  f32x16 mst[CHUNK > 0 ? 2 : 1][2][NT];
  if constexpr (CHUNK > 0) {
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) mst[p][t][ct] = f32x16{};
  }
  int nq = 0;

  // B rows l0 .. l0 + 3 start at a wave-uniform address; in a tail quad rows
  // past lmax are replaced by zeros (lanes whose row is out of range load the
  // quad's first row, which exists).
  auto load_b = [&](int l0n, float (&bo)[2][NT]) {
    const float *__restrict__ row = alm + (base + (l0n - m)) * int64_t(ncols);
    const bool full = l0n + 3 <= lmax;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const bool lok = full || (l0n + p + 2 * half <= lmax);
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) {
        const float v = row[lok ? boff[p][ct] : 0];
        bo[p][ct] = lok ? v : 0.0f;
      }
    }
  };
  // one quad: 4 recurrence steps (coefficients cq), MFMAs with operands bq
  bool check = true;
  auto quad = [&](int l0, const double2 (&cq)[4], const float (&bq)[2][NT]) {
    const bool tail = l0 + 3 > lmax;
    float f[4];
    if (check) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double pn = cq[j].x * (x * p1 - cq[j].y * p0);
        p0 = p1;
        p1 = pn;
        if (k < 0 && fabs(p1) > 0x1p300) { p1 *= 0x1p-600; p0 *= 0x1p-600; ++k; }
        f[j] = float(k == 0 ? p1 : 0.0);
      }
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double pn = cq[j].x * (x * p1 - cq[j].y * p0);
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
    if (on) {
      const auto sw0 = __builtin_amdgcn_permlane32_swap(f2u(f[0]), f2u(f[2]), false, false);
      const auto sw1 = __builtin_amdgcn_permlane32_swap(f2u(f[1]), f2u(f[3]), false, false);
      const float av[2][2] = {{u2f(sw0[0]), u2f(sw0[1])}, {u2f(sw1[0]), u2f(sw1[1])}};
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) {
          acc[p][0][ct] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[p][0], bq[p][ct], acc[p][0][ct], 0, 0, 0);
          acc[p][1][ct] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[p][1], bq[p][ct], acc[p][1][ct], 0, 0, 0);
        }
    }
    if constexpr (CHUNK > 0) {
      if (++nq == CHUNK) {
        nq = 0;
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
          for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int ct = 0; ct < NT; ++ct) {
              mst[p][t][ct] += acc[p][t][ct];
              acc[p][t][ct] = f32x16{};
            }
      }
    }
  };
  auto load_c = [&](int l0n, double2 (&cq)[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) cq[j] = co[l0n + j - m];
  };
  // two quads per iteration, operands of the next quad loaded before the
  // current one is computed (ping-pong buffers A / B, no copies)
  float bA[2][NT], bB[2][NT];
  double2 cA[4], cB[4];
  load_b(m, bA);
  load_c(m, cA);
  int l0 = m;
  for (;;) {
    const bool hasB = l0 + 4 <= lmax;
    if (hasB) { load_b(l0 + 4, bB); load_c(l0 + 4, cB); }
    quad(l0, cA, bA);
    if (!hasB) break;
    const bool hasA = l0 + 8 <= lmax;
    if (hasA) { load_b(l0 + 8, bA); load_c(l0 + 8, cA); }
    quad(l0 + 4, cB, bB);
    if (!hasA) break;
    l0 += 8;
  }
  if constexpr (CHUNK > 0) {
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) acc[p][t][ct] += mst[p][t][ct];
  }

  // epilogue: D layout of 32x32x2: lane L, acc element e -> row 8 (e/4) + 4 (L/32) + e%4,
  // column L % 32.  north = E + O, south = E - O, both times the ring phase
  // e^{i m phi0} (equal for a ring and its south mirror).  Each lane computes
  // the phase of its own ring (FP64, exact argument reduction); rows fetch it
  // by shuffle, and the (re, im) columns of a shell sit in lanes (2c, 2c+1).
  // Ring buffer layout: real column c (shell c/2) of ring r lives at
  //   ((r ngrp + c/gc)(lmax+1) + m) gc + c%gc,  ngrp = ceil(ncols / gc),
  // i.e. shells in groups of ring_group(S) = gc / 2, each (ring, group, m) row
  // 4 gc contiguous bytes.
  float phc, phs;
  {
    const int pnum = ring_pnum[rpc], pden = ring_pden[rpc];
    double ps = 0.0, pc = 1.0;
    if (pnum) sincospi(2.0 * double((int64_t(m) * pnum) % pden) / double(pden), &ps, &pc);
    phc = float(pc);
    phs = float(ps);
  }
  const int cshift = ring_group_cshift(ncols / 2), gc = 1 << cshift;
  const int ngrp = (ncols + gc - 1) >> cshift;
  const bool is_im = (col_in & 1) != 0;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int row = 32 * t + 8 * (e >> 2) + 4 * half + (e & 3);
      const float c = __shfl(phc, row, 64), sn = __shfl(phs, row, 64);
      const int rpo = rp0 + row;
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) {
        const int col = colbase + ct * 32 + col_in;
        const float ev = acc[0][t][ct][e], ov = acc[1][t][ct][e];
        const float nv = ev + ov, sv = ev - ov;
        const float no = __shfl_xor(nv, 1, 64), so = __shfl_xor(sv, 1, 64);
        // re' = re c - im s,  im' = re s + im c
        const float nrot = is_im ? (no * sn + nv * c) : (nv * c - no * sn);
        const float srot = is_im ? (so * sn + sv * c) : (sv * c - so * sn);
        if (rpo < npairs && col < ncols) {
          const float nrz = (m == 0 && is_im) ? 0.0f : nrot;
          const float srz = (m == 0 && is_im) ? 0.0f : srot;
          const int64_t gofs = ((int64_t(col >> cshift) * (lmax + 1) + m) << cshift) + (col & (gc - 1));
          F[int64_t(rpo) * ngrp * (lmax + 1) * gc + gofs] = nrz;
          if (rpo < npairs - 1)
            F[int64_t(nrings - 1 - rpo) * ngrp * (lmax + 1) * gc + gofs] = srz;
        }
      }
    }
  }
}

}  // namespace

int launch_legendre_synth_mfma(const ga_plan *p, const ga_tuning &tu, int S, const float *alm,
                               float *F, hipStream_t stream) {
  const int ncols = 2 * S;
  int nt = tu.mfma_nt;
  if (nt != 1 && nt != 2) nt = ncols <= 32 ? 1 : 2;
  const int chunk = tu.mfma_chunk;
  const int nrg = (p->npairs + 255) / 256;
  const int ncg = (ncols + 32 * nt - 1) / (32 * nt);
  const int nm8 = 4 * ((p->lmax + 1 + 31) / 32);  // m slots per XCD, runs of 4
  const int64_t nblk = int64_t(8) * nm8 * nrg * ncg;
  GA_REQUIRE(nblk < (int64_t(1) << 31), GA_ERR_UNSUPPORTED, "grid too large");
#define GA_MFMA_LAUNCH(NT_, CH_, MW_)                                                         \
  hipLaunchKernelGGL((legendre_synth_mfma_kernel<NT_, CH_, MW_>), dim3(unsigned(nblk)),       \
                     dim3(256), 0, stream, p->lmax, p->npairs, p->nrings, ncols, nrg, ncg,    \
                     p->d_z, p->d_sth, p->d_ab, p->d_cm, p->d_ring_pnum, p->d_ring_pden, alm, F)
  if (nt == 1) {
    if (chunk <= 0) GA_MFMA_LAUNCH(1, 0, 2);
    else if (chunk <= 8) GA_MFMA_LAUNCH(1, 8, 2);
    else if (chunk <= 16) GA_MFMA_LAUNCH(1, 16, 2);
    else GA_MFMA_LAUNCH(1, 32, 2);
  } else {
    if (chunk <= 0) GA_MFMA_LAUNCH(2, 0, 2);
    else if (chunk <= 8) GA_MFMA_LAUNCH(2, 8, 1);
    else if (chunk <= 16) GA_MFMA_LAUNCH(2, 16, 1);
    else GA_MFMA_LAUNCH(2, 32, 1);
  }
#undef GA_MFMA_LAUNCH
  GA_LAUNCH_CHECK();
  return GA_OK;
}

}  // namespace ga
