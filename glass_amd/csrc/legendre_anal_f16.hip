// Batched Legendre stage of the analysis (map2alm) on split-FP16 MFMA (gfx950).
//
//   a_lm[c] = sum_rp lambda_lm(theta_rp) Q_p[rp, c],   Q_0 = F_N + F_S, Q_1 = F_N - F_S,
//
// p = parity of l - m, c = real column (re / im of a shell), F the weighted,
// phase-removed ring coefficients [ring][m][c].  The sum runs over ring pairs,
// which is the K dimension of the MFMA: D[l][c] += A[l][rp] B[rp][c] with
// A = lambda (split FP16, times 2^8) and B = Q_p (split FP16, scaled per column
// and wave into [2^12, 2^13)), three v_mfma_f32_32x32x16_f16 per product as in
// the synthesis stages.
//
// Work decomposition: a workgroup = 8 waves = 512 ring pairs of one m and one
// 32-column tile; a lane owns one ring pair and runs its FP64 recurrence.  The
// l-range of m is cut into slabs of 64 rows (32 per parity, one 32-row MFMA
// tile each).  The recurrence yields lambda with ring pairs in lanes and l in
// registers, but the MFMA wants l in lanes (rows) and ring pairs in registers
// (K): each wave transposes its slab through its own LDS tile
// [hi|lo][parity][l'][pair] (16-bit stores of a lane's values, 16-byte loads
// of 8 consecutive pairs).  B (the wave's 64 pairs x 32 columns, both
// parities) stays in registers for the whole m.  After each slab the 8 waves'
// D tiles are summed in LDS in a fixed order and the workgroup writes its
// partial sum [slot][idx][c]; a second kernel adds the slots in a fixed order,
// so the result does not depend on scheduling.
#include "common.h"
#include "device.h"
#include "f16.h"
#include "kernels.h"

namespace ga {

namespace {

using f16::f32x16;
using f16::f16x8;
using f16::split_pair;

constexpr int kAWaves = 8;                // waves per workgroup
constexpr int kAPairs = 64 * kAWaves;     // ring pairs per workgroup = one partial slot
constexpr int kASlab = 64;                // l rows per slab
constexpr int kARowPad = 8;               // halves of padding per LDS row of lambda
constexpr int kAColTop = 13;              // column maxima are scaled below 2^13

__device__ __forceinline__ int anal_shift(float mx) {
  if (!(mx > 0.0f) || !isfinite(mx)) return 0;
  int e;
  (void)frexpf(mx, &e);  // mx < 2^e
  return kAColTop - e;
}

template <int NT>
__global__ __launch_bounds__(64 * kAWaves, 1) void legendre_anal_f16_kernel(
    int lmax, int npairs, int nrings, int ncols, int nrg, int ncg, const double *__restrict__ z,
    const double *__restrict__ sth, const double *__restrict__ gam,
    const double *__restrict__ csc, const double *__restrict__ cm, const float *__restrict__ F,
    float *__restrict__ partial, int64_t nalm, int lam_exp) {
  // XCD-aware decode (as in the synthesis stages): XCD x takes m = x, x + 8, ...
  const int b = blockIdx.x;
  const int xcd = b & 7, kq = b >> 3, G = nrg * ncg;
  const int m = 8 * (kq / G) + xcd;
  if (m > lmax) return;  // whole workgroup
  const int rem = kq - (kq / G) * G;
  const int rg = rem / ncg, cg = rem - (rem / ncg) * ncg;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int half = lane >> 5, c32 = lane & 31;
  const int pw0 = rg * kAPairs + wave * 64;  // first ring pair of this wave
  const int rp = pw0 + lane;
  const bool valid = rp < npairs;
  const int rpc = valid ? rp : npairs - 1;
  const double x = z[rpc];
  double p0, p1 = 0.0;
  int k;
  // the normalised recurrence of the synthesis, y_l = lambda_l / c_l:
  // y_l = (gam_l x) y_{l-1} - y_{l-2}, started from (-seed, 0); the partial
  // sums of y are multiplied by c_l when stored
  legendre_seed(ldexp(cm[m], lam_exp), sth[rpc], m, p0, k);
  p0 = -p0;
  if (!valid) {  // a missing ring pair contributes zeros
    p0 = 0.0;
    k = 0;
  }

  // B operands: lane L holds B[k = 8 (L/32) + j][n = L % 32] of k-block kb,
  // i.e. ring pair pw0 + 16 kb + 8 half + j, column col, for both parities;
  // NT 32-column tiles per wave share the recurrence values.
  uint4 bh[NT][2][4], bl[NT][2][4];
  float fac[NT];
#pragma unroll
  for (int ct = 0; ct < NT; ++ct) {
    const int col = (cg * NT + ct) * 32 + c32;
    const bool colok = col < ncols;
    float fn[4][8], fs[4][8];
    float mx = 0.0f;
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int pr = pw0 + 16 * kb + 8 * half + j;
        float vn = 0.0f, vs = 0.0f;
        if (colok && pr < npairs) {
          vn = F[(int64_t(pr) * (lmax + 1) + m) * ncols + col];
          if (pr < npairs - 1) vs = F[(int64_t(nrings - 1 - pr) * (lmax + 1) + m) * ncols + col];
        }
        fn[kb][j] = vn;
        fs[kb][j] = vs;
        mx = fmaxf(mx, fmaxf(fabsf(vn + vs), fabsf(vn - vs)));
      }
    mx = fmaxf(mx, __shfl_xor(mx, 32, 64));  // lanes L and L ^ 32 hold the same column
    const int sh = anal_shift(mx);
    fac[ct] = colok ? ldexpf(1.0f, -(lam_exp + sh)) : 0.0f;
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
      uint32_t h[2][4], l[2][4];
#pragma unroll
      for (int j = 0; j < 8; j += 2) {
        const float e0 = ldexpf(fn[kb][j] + fs[kb][j], sh), e1 = ldexpf(fn[kb][j + 1] + fs[kb][j + 1], sh);
        const float o0 = ldexpf(fn[kb][j] - fs[kb][j], sh), o1 = ldexpf(fn[kb][j + 1] - fs[kb][j + 1], sh);
        split_pair(e0, e1, h[0][j / 2], l[0][j / 2]);
        split_pair(o0, o1, h[1][j / 2], l[1][j / 2]);
      }
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        bh[ct][p][kb] = uint4{h[p][0], h[p][1], h[p][2], h[p][3]};
        bl[ct][p][kb] = uint4{l[p][0], l[p][1], l[p][2], l[p][3]};
      }
    }
    __builtin_amdgcn_sched_barrier(0);  // one tile's raw values live at a time
  }

  // LDS: per wave the transposed slab lam[hi|lo][parity][l' (32)][pair (64)],
  // rows padded by 8 halves (16 bytes) so that the 16-byte A reads of 8
  // consecutive lanes (rows 144 bytes apart) fall in distinct banks (18 KB);
  // after the MFMAs of a slab the same memory holds the waves' D tiles for
  // the fixed-order sum.  Recurrence coefficients gam_l are the same for
  // every wave of the workgroup (one m): one double buffer of one slab,
  // written by wave 0, read after the slab's barriers.
  __shared__ __align__(16) _Float16 lam[kAWaves][2][2][32][64 + kARowPad];
  __shared__ double cbuf[2][kASlab];
  float *red = reinterpret_cast<float *>(&lam[0][0][0][0][0]);  // [wave][p][e][lane]

  const int64_t base = hp_index(lmax, m, m);
  const double *__restrict__ co = gam + base;
  const int lastc = lmax - m;
  double cpre = co[min(lane, lastc)];
  if (wave == 0) cbuf[0][lane] = cpre;
  __syncthreads();
  _Float16(*lw)[2][32][64 + kARowPad] = lam[wave];  // [hi|lo][p][l'][pair]

  const int nsl = (lmax - m + kASlab) / kASlab;
  // Leading slabs in which every lane of the workgroup either starts scaled
  // (k < 0: values below 2^-300) or keeps |y 2^lam_exp| below 2^-25, where both
  // FP16 halves of the split are exactly 0, give partial sums of exactly 0:
  // only the recurrence advances (one rescaling check per slab; power-of-two
  // rescaling is exact, so the state matches the per-group checks bit for
  // bit) and the slab's rows of the partial are written as 0.  The slab in
  // which some lane reaches FP16 range is undone and run by the full loop.
  int s0 = 0;
  while (s0 < nsl && m + kASlab * (s0 + 1) - 1 <= lmax) {
    const int cur = s0 & 1;
    cpre = co[min(kASlab * (s0 + 1) + lane, lastc)];
    const double q0 = p0, q1 = p1;
    const int kq = k;
    double mx = 0.0;
#pragma unroll 8
    for (int j = 0; j < kASlab; ++j) {
      const double c = cbuf[cur][j];
      const double pn = fma(c * x, p1, -p0);
      p0 = p1;
      p1 = pn;
      mx = fmax(mx, fabs(pn));
    }
    if (k < 0 && fabs(p1) > 0x1p300) {
      p1 *= 0x1p-600;
      p0 *= 0x1p-600;
      ++k;
    }
    if (wave == 0) cbuf[cur ^ 1][lane] = cpre;  // read after this slab's barriers
    if (!__syncthreads_and(kq < 0 || mx < 0x1p-25)) {  // values in this slab: undo it
      p0 = q0;
      p1 = q1;
      k = kq;
      break;
    }
    const int l0 = m + kASlab * s0;
    for (int e = threadIdx.x; e < kASlab * 32 * NT; e += 64 * kAWaves) {
      const int l = l0 + e / (32 * NT);
      const int col = cg * NT * 32 + e % (32 * NT);
      if (col < ncols) partial[(int64_t(rg) * nalm + base + (l - m)) * ncols + col] = 0.0f;
    }
    ++s0;
  }
  for (int s = s0; s < nsl; ++s) {
    const int l0 = m + kASlab * s;
    const int cur = s & 1;
    cpre = co[min(kASlab * (s + 1) + lane, lastc)];
    const double *__restrict__ cb = cbuf[cur];
    const bool slow = (l0 + kASlab - 1 > lmax) || wave_any(k < 0);
    // 16 groups of 4 steps: step i -> parity i & 1, row l' = i >> 1
#pragma unroll 2
    for (int g = 0; g < 16; ++g) {
      float v[4];
      if (!slow) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const double c = cb[4 * g + j];
          const double pn = fma(c * x, p1, -p0);
          p0 = p1;
          p1 = pn;
          v[j] = float(pn);
        }
      } else {
        const int lg = l0 + 4 * g;
        double d[4] = {0.0, 0.0, 0.0, 0.0};
        if (lg <= lmax) {
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const double c = cb[4 * g + j];
            const double pn = fma(c * x, p1, -p0);
            p0 = p1;
            p1 = pn;
            d[j] = pn;
          }
          if (k != 0) {
#pragma unroll
            for (int j = 0; j < 4; ++j) d[j] = 0.0;
          }
          if (k < 0 && fabs(p1) > 0x1p300) {
            p1 *= 0x1p-600;
            p0 *= 0x1p-600;
            ++k;
          }
          // rows past lmax are computed but never stored: no masking needed
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = float(d[j]);
      }
      uint32_t h0, l0w, h1, l1w;
      split_pair(v[0], v[2], h0, l0w);  // parity 0, rows 2g, 2g + 1
      split_pair(v[1], v[3], h1, l1w);  // parity 1
      const f16::f16x2 H0 = __builtin_bit_cast(f16::f16x2, h0), L0 = __builtin_bit_cast(f16::f16x2, l0w);
      const f16::f16x2 H1 = __builtin_bit_cast(f16::f16x2, h1), L1 = __builtin_bit_cast(f16::f16x2, l1w);
      lw[0][0][2 * g][lane] = H0[0];
      lw[0][0][2 * g + 1][lane] = H0[1];
      lw[1][0][2 * g][lane] = L0[0];
      lw[1][0][2 * g + 1][lane] = L0[1];
      lw[0][1][2 * g][lane] = H1[0];
      lw[0][1][2 * g + 1][lane] = H1[1];
      lw[1][1][2 * g][lane] = L1[0];
      lw[1][1][2 * g + 1][lane] = L1[1];
    }
    if (wave == 0) cbuf[cur ^ 1][lane] = cpre;  // read after this slab's barriers

    // MFMAs: A[l'][pair] of k-block kb: lane L reads row L % 32, pairs
    // 16 kb + 8 (L / 32) .. + 7 (16 bytes); one A read serves the NT tiles
    f32x16 acc[NT][2];
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) acc[ct][0] = acc[ct][1] = f32x16{};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const f16x8 ah = *reinterpret_cast<const f16x8 *>(&lw[0][p][c32][16 * kb + 8 * half]);
        const f16x8 al = *reinterpret_cast<const f16x8 *>(&lw[1][p][c32][16 * kb + 8 * half]);
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) {
          const f16x8 Bh = __builtin_bit_cast(f16x8, bh[ct][p][kb]);
          const f16x8 Bl = __builtin_bit_cast(f16x8, bl[ct][p][kb]);
          acc[ct][p] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, Bh, acc[ct][p], 0, 0, 0);
          acc[ct][p] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, Bl, acc[ct][p], 0, 0, 0);
          acc[ct][p] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, Bh, acc[ct][p], 0, 0, 0);
        }
      }
    }

    // fixed-order sum of the 8 waves' tiles: element e of lane L is row
    // 8 (e/4) + 4 (L/32) + e%4 (l = l0 + 2 row + p), column L % 32 of tile ct
    __syncthreads();  // every wave has read its lambda tile
#pragma unroll
    for (int ct = 0; ct < NT; ++ct)
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int e = 0; e < 16; ++e)
          red[(((wave * NT + ct) * 2 + p) * 16 + e) * 64 + lane] = acc[ct][p][e] * fac[ct];
    __syncthreads();
#pragma unroll 2
    for (int q = 0; q < 4 * NT; ++q) {
      const int pe = wave * 4 * NT + q;  // (tile, parity, element) handled by this wave
      const int ct = pe >> 5, p = (pe >> 4) & 1, e = pe & 15;
      float sum = 0.0f;
#pragma unroll
      for (int w = 0; w < kAWaves; ++w) sum += red[(((w * NT + ct) * 2 + p) * 16 + e) * 64 + lane];
      const int row = 8 * (e >> 2) + 4 * half + (e & 3);
      const int l = l0 + 2 * row + p;
      const int col = (cg * NT + ct) * 32 + c32;
      if (l <= lmax && col < ncols)
        partial[(int64_t(rg) * nalm + base + (l - m)) * ncols + col] =
            (m == 0 && (col & 1)) ? 0.0f : float(double(sum) * csc[base + (l - m)]);
    }
    __syncthreads();  // the sums are read before the next slab's lambda tile
  }
}

}  // namespace

int launch_legendre_anal_f16(const ga_plan *p, int S, const float *F, float *partial,
                             hipStream_t stream) {
  const int ncols = 2 * S;
  const int nrg = (p->npairs + kAPairs - 1) / kAPairs;
  GA_REQUIRE(nrg == anal_partial_slots(p), GA_ERR_UNSUPPORTED, "partial slots %d != %d", nrg,
             anal_partial_slots(p));
  // two 32-column tiles per wave when there are more than 32 columns (the
  // recurrence then serves 64 columns; the waves' tiles fill the reduction
  // buffer exactly)
  const int nt = ncols > 32 ? 2 : 1;
  const int ncg = (ncols + 32 * nt - 1) / (32 * nt);
  const int nm8 = (p->lmax + 1 + 7) / 8;
  const int64_t nblk = int64_t(8) * nm8 * nrg * ncg;
  GA_REQUIRE(nblk < (int64_t(1) << 31), GA_ERR_UNSUPPORTED, "grid too large");
  if (nt == 2)
    hipLaunchKernelGGL(legendre_anal_f16_kernel<2>, dim3(unsigned(nblk)), dim3(64 * kAWaves), 0,
                       stream, p->lmax, p->npairs, p->nrings, ncols, nrg, ncg, p->d_z, p->d_sth,
                       p->d_gam, p->d_csc, p->d_cm, F, partial, p->nalm, p->lam_exp16);
  else
    hipLaunchKernelGGL(legendre_anal_f16_kernel<1>, dim3(unsigned(nblk)), dim3(64 * kAWaves), 0,
                       stream, p->lmax, p->npairs, p->nrings, ncols, nrg, ncg, p->d_z, p->d_sth,
                       p->d_gam, p->d_csc, p->d_cm, F, partial, p->nalm, p->lam_exp16);
  GA_LAUNCH_CHECK();
  return GA_OK;
}

}  // namespace ga
