// Spin-s synthesis Legendre stage on split-FP16 MFMA (gfx950), s = 1, 2.
//
// HEALPix convention Q + iU = -sum (E + iB) sY_lm.  For E-mode coefficients
// (B = 0) the per-ring Fourier coefficients are
//     Q_m = sum_l E_lm F+_lm(theta),    U_m = -i sum_l E_lm F-_lm(theta),
//     F+- = -(lambda+ +- (-1)^s lambda-) / 2,
// with lambda+- the normalised Wigner-d functions of spin +s and -s (the
// three-term recurrence of the plan's spin tables, in FP64 with the scaled
// seed of the scalar stage).  The contraction is linear in lambda, so the
// kernel contracts lambda' = -2^7 lambda+ and mu = (-1)^s (-2^7) lambda-
// separately and combines the accumulators afterwards:
//     2^8 F+ = acc(lambda') + acc(mu),   2^8 F- = acc(lambda') - acc(mu).
// A wave owns 32 ring pairs: lanes 0..31 run the +s recurrence of ring pair
// L, lanes 32..63 the -s recurrence of ring pair L - 32.  The permlane swap
// of the scalar stage then yields the lambda' tile (row tile 0) and the mu
// tile (row tile 1) of the same 32 rings, both against the same E operand:
// per (ring pair, l), two FP64 recurrence steps and twice the MFMAs of the
// scalar stage, and no cross-lane traffic before the epilogue.
// South rings: F+(pi - theta) = sigma F+(theta), F-(pi - theta) = -sigma F-(theta),
// sigma = (-1)^(l+m+s); with E / O the even / odd l - m parts,
//     north  P+ = E+ + O+,            P- = E- + O-,
//     south  P+ = (-1)^s (E+ - O+),   P- = -(-1)^s (E- - O-),
// Q = P+, U = -i P-.  At m = 0 the imaginary part of the coefficients is
// ignored (Q_0 real, U_0 imaginary), as in the FP64 stage.
// B != 0 reuses the kernel: the maps of (0, B) are (-U, Q) of the maps of
// (B, 0), so a second pass over B adds (-U_B, Q_B) into the ring buffer.
//
// Output: the grouped ring buffer of the FP32 ring stage for 2S transforms,
// transform 2s = Q_s and 2s + 1 = U_s, with the ring phase e^{i m phi0} applied.
#include "common.h"
#include "device.h"
#include "f16.h"
#include "kernels.h"

namespace ga {

namespace {

using f16::f32x16;
using f16::kLamExp;
using f16::split_pair;

constexpr int kSpinPairs = 32;  // ring pairs per wave

template <int NT, int MINW>
__global__ __launch_bounds__(256, MINW) void spin_synth_f16_kernel(
    int lmax, int spin, int npairs, int nrings, int ncols, int nct, int nrg, int ncg,
    const double *__restrict__ z, const double *__restrict__ sth,
    const double *__restrict__ tab /* (A, A B, C) per (l, m) */,
    const double2 *__restrict__ seeds /* [+s | -s][m]: (signed mantissa, exponent) */,
    const int32_t *__restrict__ ring_pnum, const int32_t *__restrict__ ring_pden,
    const int64_t *__restrict__ slab_off, const uint4 *__restrict__ B,
    const int32_t *__restrict__ colexp, float *__restrict__ F, int bpass) {
  // XCD-aware decode as in the scalar stage: XCD x takes the runs of four
  // consecutive m starting at 4 x, 4 x + 32, ... (the four m of a 128-byte
  // line of the ring buffer are written through one XCD's L2), all (ring
  // group, column group) tiles of one m back to back.
  const int b = blockIdx.x;
  const int xcd = b & 7, kq = b >> 3, G = nrg * ncg;
  const int mq = kq / G;
  const int m = 32 * (mq >> 2) + 4 * xcd + (mq & 3);
  if (m > lmax) return;
  const int rem = kq - mq * G;
  const int rg = rem / ncg, cg = rem - (rem / ncg) * ncg;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int half = lane >> 5, col_in = lane & 31;
  const int rp0 = rg * (4 * kSpinPairs) + wave * kSpinPairs;
  if (rp0 >= npairs) return;  // wave-uniform; the kernel has no barriers
  const int rp = rp0 + col_in;
  const int rpc = rp < npairs ? rp : npairs - 1;
  const double x = z[rpc];

  // seed of this lane's recurrence at l0 = max(m, s), placed so that the
  // (0, 0, -1) steps for l = m .. l0 deliver it at l0
  double p0, p1;
  int k;
  {
    const double st = sth[rpc];
    double sh = sqrt(0.5 * (1.0 - x)), ch = sqrt(0.5 * (1.0 + x));
    if (x > 0.5) sh = st / (2.0 * ch);
    else if (x < -0.5) ch = st / (2.0 * sh);
    const int dif = m > spin ? m - spin : spin - m;
    const int ns = half ? dif : m + spin, nc = half ? m + spin : dif;
    double msn, mcs;
    long long es, ec;
    scaled_power(sh, ns, msn, es);
    scaled_power(ch, nc, mcs, ec);
    const double2 kk = seeds[half * (lmax + 1) + m];
    // lambda' = -2^7 lambda+, mu = -(-1)^s 2^7 lambda-
    const double sg = (half && (spin & 1)) ? 1.0 : -1.0;
    int t;
    const double v = frexp(sg * kk.x * msn * mcs, &t);
    double sd;
    to_scaled(v, (long long)kk.y + es + ec + t + (kLamExp - 1), sd, k);
    const int l0 = m > spin ? m : spin;
    if ((l0 - m) & 1) {
      p0 = 0.0;
      p1 = sd;
    } else {
      p0 = sd;
      p1 = 0.0;
    }
  }

  const int64_t base = hp_index(lmax, m, m);
  const double *__restrict__ co = tab + 3 * base;
  const int64_t s0 = slab_off[m];
  const int nsl = int(nslab_of(lmax, m));
  const int colbase = cg * 32 * NT;
  int ctile[NT];
#pragma unroll
  for (int ct = 0; ct < NT; ++ct) ctile[ct] = min(cg * NT + ct, nct - 1);

  f32x16 acc[2][2][NT];  // [parity][lambda' | mu][col tile]
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) acc[p][t][ct] = f32x16{};

  // Recurrence coefficients through LDS, double-buffered per wave (as in the
  // scalar stage): (A, +-A B) per half-wave and C; lane L loads row L % 32.
  __shared__ double2 cu[4][2][2][kSlab];
  __shared__ double cw[4][2][kSlab];
  const int lastc = lmax - m;
  const double hs = half ? -1.0 : 1.0;
  double fA, fB, fC;
  {
    const double *__restrict__ q = co + 3 * min(col_in, lastc);
    fA = q[0];
    fB = q[1];
    fC = q[2];
  }
  cu[wave][0][half][col_in] = double2{fA, hs * fB};
  if (!half) cw[wave][0][col_in] = fC;
  __builtin_amdgcn_wave_barrier();

  for (int s = 0; s < nsl; ++s) {
    const int l0 = m + kSlab * s;
    const int cur = s & 1;
    uint4 bh[2][NT], bl[2][NT];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) {
        const uint4 *__restrict__ u = B + (((s0 + s) * 2 + p) * nct + ctile[ct]) * 128;
        bh[p][ct] = u[lane];
        bl[p][ct] = u[64 + lane];
      }
    {
      const double *__restrict__ q = co + 3 * min(kSlab * (s + 1) + col_in, lastc);
      fA = q[0];
      fB = q[1];
      fC = q[2];
    }
    const double2 *__restrict__ cb = cu[wave][cur][half];
    const double *__restrict__ cc = cw[wave][cur];

    uint32_t ph[2][8], pl[2][8];
    const bool slow = (l0 + kSlab - 1 > lmax) || wave_any(k < 0);
    if (!slow) {
#pragma unroll
      for (int g = 0; g < 8; ++g) {
        float v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const double2 u = cb[4 * g + j];
          const double pn = fma(fma(u.x, x, u.y), p1, -(cc[4 * g + j] * p0));
          p0 = p1;
          p1 = pn;
          v[j] = float(pn);
        }
        split_pair(v[0], v[2], ph[0][g], pl[0][g]);
        split_pair(v[1], v[3], ph[1][g], pl[1][g]);
        __builtin_amdgcn_sched_barrier(0);
      }
    } else {
#pragma unroll
      for (int g = 0; g < 8; ++g) {
        const int lg = l0 + 4 * g;
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        bool dead = true;
        if (lg <= lmax) {
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const double2 u = cb[4 * g + j];
            const double pn = fma(fma(u.x, x, u.y), p1, -(cc[4 * g + j] * p0));
            p0 = p1;
            p1 = pn;
            v[j] = p1;
          }
          // as in the scalar stage: a scaled lane contributes zero; rescale
          // once per group
          dead = wave_all(k < 0);
          if (k != 0) {
#pragma unroll
            for (int j = 0; j < 4; ++j) v[j] = 0.0;
          }
          if (k < 0 && fabs(p1) > 0x1p300) {
            p1 *= 0x1p-600;
            p0 *= 0x1p-600;
            ++k;
          }
          if (lg + 3 > lmax) {
#pragma unroll
            for (int j = 0; j < 4; ++j)
              if (lg + j > lmax) v[j] = 0.0;
          }
        }
        if (dead) {
          ph[0][g] = ph[1][g] = pl[0][g] = pl[1][g] = 0u;
        } else {
          split_pair(float(v[0]), float(v[2]), ph[0][g], pl[0][g]);
          split_pair(float(v[1]), float(v[3]), ph[1][g], pl[1][g]);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    cu[wave][cur ^ 1][half][col_in] = double2{fA, hs * fB};
    if (!half) cw[wave][cur ^ 1][col_in] = fC;
    __builtin_amdgcn_wave_barrier();

    if (wave_any(k == 0)) f16::slab_mfma<NT>(ph, pl, bh, bl, acc);
  }

  // epilogue: element e of lane L is ring row 8 (e/4) + 4 (L/32) + e%4 of the
  // wave's 32 ring pairs, real column L % 32 (re / im of a shell in lanes 2c,
  // 2c + 1).  Unscale, combine, rotate by the ring phase and store Q and U.
  float phc, phs;
  {
    const int pnum = ring_pnum[rpc], pden = ring_pden[rpc];
    double ps = 0.0, pc = 1.0;
    if (pnum) sincospi(2.0 * double((int64_t(m) * pnum) % pden) / double(pden), &ps, &pc);
    phc = float(pc);
    phs = float(ps);
  }
  float fac[NT];
#pragma unroll
  for (int ct = 0; ct < NT; ++ct) {
    const int col = colbase + ct * 32 + col_in;
    fac[ct] = col < ncols ? ldexpf(1.0f, -(kLamExp + colexp[col])) : 0.0f;
  }
  // 2S transforms = 4S real columns, in groups of ring_group(2S)
  const int cshift = ring_group_cshift(ncols), gc = 1 << cshift;
  const int ngrp = (2 * ncols + gc - 1) >> cshift;
  const int64_t ring_stride = int64_t(ngrp) * (lmax + 1) * gc;
  const bool is_im = (col_in & 1) != 0;
  const float ss = (spin & 1) ? -1.0f : 1.0f;
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int row = 8 * (e >> 2) + 4 * half + (e & 3);
    const float c = __shfl(phc, row, 64), sn = __shfl(phs, row, 64);
    const int rpo = rp0 + row;
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) {
      const int col = colbase + ct * 32 + col_in;
      const float f = fac[ct];
      const float a0 = acc[0][0][ct][e] * f, a1 = acc[0][1][ct][e] * f;
      const float b0 = acc[1][0][ct][e] * f, b1 = acc[1][1][ct][e] * f;
      const float ep = a0 + a1, em = a0 - a1, op = b0 + b1, om = b0 - b1;
      float P[2] = {ep + op, ss * (ep - op)};    // north, south: this lane's part of P+
      float M[2] = {em + om, -ss * (em - om)};  // and of P-
      float qv[2], uv[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const float po = __shfl_xor(P[h], 1, 64), mo = __shfl_xor(M[h], 1, 64);
        // Q = P+ and U = -i P- = (P-_im, -P-_re), each times (c + i sn)
        qv[h] = is_im ? (po * sn + P[h] * c) : (P[h] * c - po * sn);
        uv[h] = is_im ? (M[h] * sn - mo * c) : (mo * c + M[h] * sn);
        if (m == 0) {
          if (is_im) qv[h] = 0.0f;
          else uv[h] = 0.0f;
        }
      }
      if (rpo < npairs && col < ncols) {
        const int qc = 4 * (col >> 1) + (col & 1), uc = qc + 2;
        const int64_t oq = ((int64_t(qc >> cshift) * (lmax + 1) + m) << cshift) + (qc & (gc - 1));
        const int64_t ou = ((int64_t(uc >> cshift) * (lmax + 1) + m) << cshift) + (uc & (gc - 1));
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          if (h == 1 && rpo >= npairs - 1) break;  // the equator ring has no mirror
          const int64_t rb = int64_t(h ? nrings - 1 - rpo : rpo) * ring_stride;
          if (bpass) {
            F[rb + oq] -= uv[h];
            F[rb + ou] += qv[h];
          } else {
            F[rb + oq] = qv[h];
            F[rb + ou] = uv[h];
          }
        }
      }
    }
  }
}

}  // namespace

int launch_spin_synth_f16(const ga_plan *p, int spin, int S, const float *E, const float *Bm,
                          void *ws, hipStream_t stream) {
  GA_REQUIRE(spin >= 1 && spin <= 2 && p->d_spin[spin] && p->d_spin_seed[spin], GA_ERR_INVALID,
             "spin %d tables not prepared (call ga_plan_prepare_spin)", spin);
  const int ncols = 2 * S;
  const int nct = (ncols + 31) / 32;
  const f32_ws_layout w = f32_workspace(p, 2 * S);
  char *base = static_cast<char *>(ws);
  float *F = reinterpret_cast<float *>(base);
  void *bws = base + w.b_off;
  int32_t *colexp = reinterpret_cast<int32_t *>(base + w.e_off);
  // two 32-column tiles per wave when there are more than 32 columns (each
  // recurrence then serves 64 columns); the fast path reads its coefficients
  // from LDS step by step, which keeps the two-tile form within the 256 VGPRs
  // of two waves per SIMD
  const int nt = ncols > 32 ? 2 : 1;
  const int nrg = (p->npairs + 4 * kSpinPairs - 1) / (4 * kSpinPairs);
  const int ncg = (ncols + 32 * nt - 1) / (32 * nt);
  const int nm8 = 4 * ((p->lmax + 1 + 31) / 32);  // m slots per XCD, runs of 4
  const int64_t nblk = int64_t(8) * nm8 * nrg * ncg;
  GA_REQUIRE(nblk < (int64_t(1) << 31), GA_ERR_UNSUPPORTED, "grid too large");
  for (int pass = 0; pass < (Bm ? 2 : 1); ++pass) {
    int rc = launch_legendre_prepare_f16(p, S, pass ? Bm : E, bws, colexp, stream, spin);
    if (rc) return rc;
#define GA_SPIN_LAUNCH(NT_, MW_)                                                                 \
  hipLaunchKernelGGL((spin_synth_f16_kernel<NT_, MW_>), dim3(unsigned(nblk)), dim3(256), 0,     \
                     stream, p->lmax, spin, p->npairs, p->nrings, ncols, nct, nrg, ncg, p->d_z, \
                     p->d_sth, p->d_spin[spin], p->d_spin_seed[spin], p->d_ring_pnum,            \
                     p->d_ring_pden, p->d_slab_off, static_cast<const uint4 *>(bws), colexp, F,  \
                     pass)
    if (nt == 2) GA_SPIN_LAUNCH(2, 2);
    else GA_SPIN_LAUNCH(1, 2);
#undef GA_SPIN_LAUNCH
    GA_LAUNCH_CHECK();
  }
  return GA_OK;
}

}  // namespace ga
