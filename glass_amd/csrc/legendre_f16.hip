// Batched Legendre stage of the scalar synthesis on split-FP16 MFMA (gfx950).
//
// Same contraction as the FP32-MFMA stage,
//     F_m[ring, s] = sum_l lambda_lm(theta_ring) a_lm[s],
// split by the parity of l - m (north = E + O, south = E - O), but each FP32
// operand is carried as a pair of FP16 numbers, x = x_hi + x_lo (22 significant
// bits), and the product is formed from three v_mfma_f32_32x32x16_f16:
//     lambda a ~ lambda_hi a_hi + lambda_hi a_lo + lambda_lo a_hi
// (the dropped lambda_lo a_lo is 2^-22 relative).  Every MFMA sums 16 products
// before adding to the FP32 accumulator, so the accumulator is rounded once per
// 16 terms instead of once per 2 (FP32 MFMA): the split path is both faster
// (FP16 MFMA has 16x the FP32 rate) and no less accurate.
//
// The recurrence runs on y_l = lambda_l / c_l (c_m = c_{m+1} = 1,
// c_l = a_l b_l c_{l-2}), where the lambda_{l-2} coefficient is 1:
// y_l = (gam_l x) y_{l-1} - y_{l-2}, one FP64 multiply and one FMA per step
// instead of two multiplies and an FMA; the B operand carries c_l a_lm, so
// sum_l y_l (c_l a_lm) = sum_l lambda_l a_lm.  c_l stays within [0.14, 1.13]
// for lmax <= 8192.
//
// Ranges: the normalised lambda_lm is bounded by sqrt((2l+1)/4pi), so
// |y_l| <= max_l sqrt((2l+1)/4pi) / c_l (151 at lmax 4096, 255 at 8192); the
// recurrence carries y 2^e with the largest e <= 8 that keeps this bound
// below 2^15.5 (e = 8 up to lmax 4096), so that it fits FP16 and small values
// keep their low part in the normal range.  The a_lm of each real column are
// scaled by a power of two that puts the column's maximum in [2^12, 2^13)
// (times c_l <= 1.13 after the c scaling).  The accumulator is unscaled in the
// epilogue.
//
// Work decomposition as in the FP32 stage: a workgroup = 4 waves = 256 ring
// pairs of one m and NT 32-column tiles; a lane owns one ring pair and runs the
// FP64 recurrence (scaled seed, no underflow).  The l-range of m is cut into
// slabs of 32 rows (K = 16 per parity).  Per slab a lane forms its 32 lambda
// values, splits and packs them as FP16 pairs, and v_permlane32_swap moves the
// upper K half between the wave halves so that each 32-ring row tile holds the
// 32 x 16 A operand (lane L: ring L % 32, k = 8 (L / 32) .. + 7).  B comes
// pre-split and pre-swizzled (one 16-byte load per lane for the high and one
// for the low part), prepared once per batch by the two small kernels below.
#include <type_traits>

#include "common.h"
#include "device.h"
#include "f16.h"
#include "kernels.h"

namespace ga {

namespace {

using f16::f32x16;
using f16::split_pair;

constexpr int kColTop = 13;  // column maxima are scaled below 2^13


__global__ void fill_u32_kernel(uint32_t *__restrict__ p, int n, uint32_t v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// per-column max |a| as float bit patterns (non-negative floats order like
// unsigned integers); 64 columns x 4 row streams per block
__global__ __launch_bounds__(256) void colmax_kernel(int64_t nrows, int ncols,
                                                     const float *__restrict__ alm,
                                                     uint32_t *__restrict__ cmax) {
  __shared__ uint32_t red[4][64];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int c0 = 0; c0 < ncols; c0 += 64) {
    const int c = c0 + tx;
    uint32_t mx = 0;
    if (c < ncols)
      for (int64_t r = int64_t(blockIdx.x) * 4 + ty; r < nrows; r += int64_t(gridDim.x) * 4)
        mx = max(mx, __float_as_uint(fabsf(alm[r * ncols + c])));
    red[ty][tx] = mx;
    __syncthreads();
    if (ty == 0 && c < ncols) {
      mx = max(max(red[0][tx], red[1][tx]), max(red[2][tx], red[3][tx]));
      if (mx) atomicMax(&cmax[c], mx);
    }
    __syncthreads();
  }
}

// the same for ncols % 4 == 0 and ncols <= 1024: each thread reads float4s of
// four fixed columns (T4 = ncols / 4 threads per row, 256 / T4 rows per pass)
__global__ __launch_bounds__(256) void colmax4_kernel(int64_t nrows, int ncols,
                                                      const float4 *__restrict__ alm,
                                                      uint32_t *__restrict__ cmax) {
  __shared__ uint32_t red[4][256];
  const int T4 = ncols >> 2, R = 256 / T4;
  const int tid = threadIdx.x;
  const int r = tid / T4, q = tid - r * T4;
  uint32_t mx[4] = {0u, 0u, 0u, 0u};
  if (r < R) {
    for (int64_t row = int64_t(blockIdx.x) * R + r; row < nrows; row += int64_t(gridDim.x) * R) {
      const float4 v = alm[row * T4 + q];
      mx[0] = max(mx[0], __float_as_uint(fabsf(v.x)));
      mx[1] = max(mx[1], __float_as_uint(fabsf(v.y)));
      mx[2] = max(mx[2], __float_as_uint(fabsf(v.z)));
      mx[3] = max(mx[3], __float_as_uint(fabsf(v.w)));
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) red[j][tid] = mx[j];
  __syncthreads();
  if (tid < T4) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      uint32_t m = red[j][tid];
      for (int rr = 1; rr < R; ++rr) m = max(m, red[j][rr * T4 + tid]);
      if (m) atomicMax(&cmax[4 * tid + j], m);
    }
  }
}

__device__ __forceinline__ int col_shift(uint32_t bits) {
  const float mx = __uint_as_float(bits);
  if (!(mx > 0.0f) || !isfinite(mx)) return 0;
  int e;
  (void)frexpf(mx, &e);  // mx < 2^e
  return kColTop - e;
}

// B operand units: unit = (slab * 2 + parity) * nct + column tile; lane L of a
// unit holds rows l0 + p + 16 (L / 32) + 2 j (j = 0..7) of column tile * 32 +
// L % 32, as 8 high halves (first 64 x 16 bytes) and 8 low halves (next).
// grid: (ceil(units of m = 0 / 4), lmax + 1): blockIdx.y = m, 4 units per block.
// Rows l < lmin are stored as zero (the spin-s stages: no modes below l = s).
// With csc, row l of m is multiplied by csc[(m, l)] (in FP64, rounded once)
// before the scaling: the c_l of the normalised synthesis recurrence.
__global__ __launch_bounds__(256) void swizzle_b_kernel(int lmax, int lmin, int ncols, int nct,
                                                        const int64_t *__restrict__ slab_off,
                                                        const float *__restrict__ alm,
                                                        const uint32_t *__restrict__ cmax,
                                                        const double *__restrict__ csc,
                                                        uint4 *__restrict__ B,
                                                        int32_t *__restrict__ colexp) {
  const int m = blockIdx.y;
  const int u = blockIdx.x * 4 + (threadIdx.x >> 6);  // unit within m
  const int L = threadIdx.x & 63;
  const int per_slab = 2 * nct;
  if (u >= int(nslab_of(lmax, m)) * per_slab) return;
  const int sl = u / per_slab, rem = u - sl * per_slab;
  const int p = rem / nct, ct = rem - p * nct;
  const int64_t sg = slab_off[m] + sl;
  const int64_t unit = sg * per_slab + rem;
  const int l0 = m + kSlab * sl;
  const int c = ct * 32 + (L & 31);
  const bool cok = c < ncols;
  const int cc = cok ? c : 0;
  // all eight loads are issued before any is used (rows outside [lmin, lmax]
  // and padding columns read the in-range element (m, cc) and are zeroed)
  const int64_t mbase = hp_index(lmax, m, m);
  const float *__restrict__ col = alm + mbase * int64_t(ncols) + cc;
  float raw[8];
  bool ok[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int l = l0 + p + 16 * (L >> 5) + 2 * j;
    ok[j] = cok && l <= lmax && l >= lmin;
    raw[j] = col[int64_t(ok[j] ? l - m : 0) * ncols];
  }
  if (csc) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int l = l0 + p + 16 * (L >> 5) + 2 * j;
      raw[j] = float(double(raw[j]) * csc[mbase + (ok[j] ? l - m : 0)]);
    }
  }
  const int sh = cok ? col_shift(cmax[cc]) : 0;
  if (sg == 0 && p == 0 && L < 32) colexp[c] = sh;
  uint32_t hv[4], lv[4];
#pragma unroll
  for (int j = 0; j < 8; j += 2) {
    const float v0 = ok[j] ? ldexpf(raw[j], sh) : 0.0f;
    const float v1 = ok[j + 1] ? ldexpf(raw[j + 1], sh) : 0.0f;
    split_pair(v0, v1, hv[j / 2], lv[j / 2]);
  }
  B[unit * 128 + L] = uint4{hv[0], hv[1], hv[2], hv[3]};
  B[unit * 128 + 64 + L] = uint4{lv[0], lv[1], lv[2], lv[3]};
}

// One slab (32 rows from l0) of a lane's recurrence: y_l = lambda_l / c_l
// times 2^lam_exp for its ring pair (y_l = (gam_l x) y_{l-1} - y_{l-2}; the
// B operand carries c_l a_lm), split and packed as FP16 pairs (ph / pl
// [parity][group], group g = rows k = 2g, 2g + 1 of each parity), with the
// scaled-seed bookkeeping of the turn-on region.  cb: the slab's gam_l in LDS
// (16-byte aligned).
__device__ __forceinline__ void slab_values(const double *__restrict__ cb, double x, int l0,
                                            int lmax, double &p0, double &p1, int &k,
                                            uint32_t (&ph)[2][8], uint32_t (&pl)[2][8]) {
  const bool slow = (l0 + kSlab - 1 > lmax) || wave_any(k < 0);
  if (!slow) {
    // every lane in range and a full slab: the plain recurrence.  The
    // coefficients of group g + 1 are read from LDS while group g runs; the
    // scheduling barrier keeps later groups' reads from being hoisted
    // further (each would hold 8 VGPRs for the rest of the slab).
    const double2 *__restrict__ cb2 = reinterpret_cast<const double2 *>(cb);
    double2 cc[2], cn[2];
    cc[0] = cb2[0];
    cc[1] = cb2[1];
#pragma unroll
    for (int g = 0; g < 8; ++g) {
      if (g < 7) {
        cn[0] = cb2[2 * (g + 1)];
        cn[1] = cb2[2 * (g + 1) + 1];
      }
      const double gm[4] = {cc[0].x, cc[0].y, cc[1].x, cc[1].y};
      float v[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double pn = fma(gm[j] * x, p1, -p0);
        p0 = p1;
        p1 = pn;
        v[j] = float(pn);
      }
      split_pair(v[0], v[2], ph[0][g], pl[0][g]);
      split_pair(v[1], v[3], ph[1][g], pl[1][g]);
      cc[0] = cn[0];
      cc[1] = cn[1];
      __builtin_amdgcn_sched_barrier(0);
    }
  } else {
#pragma unroll
    for (int g = 0; g < 8; ++g) {
      const int lg = l0 + 4 * g;
      double v[4] = {0.0, 0.0, 0.0, 0.0};
      bool dead = true;  // every lane still scaled: the group's values are all zero
      if (lg <= lmax) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const double c = cb[4 * g + j];
          const double pn = fma(c * x, p1, -p0);
          p0 = p1;
          p1 = pn;
          v[j] = p1;
        }
        // k is fixed within a group: a scaled lane's values are < 2^-290
        // (zero in FP32); rescale once per group (4 steps grow |p| by far
        // less than the 2^700 of headroom above the threshold)
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
}

// The dead leading slabs of a wave: only the recurrence advances, without
// values, splitting, B loads or MFMAs, and with one rescaling check per slab
// (a slab grows |lambda| by far less than the 2^700 above the threshold;
// power-of-two rescaling is exact, so the state matches the per-group checks
// of slab_values bit for bit).  The slab in which some lane reaches FP16
// range is undone, to be run again in full; returns its index (the first
// slab to run in full).  A slab is dead when every lane either starts it
// scaled (k < 0: all its values are below 2^-300) or keeps its carried value
// (y 2^lam_exp) below 2^-25 in magnitude throughout, where both FP16 halves of the split are exactly 0 (the
// high half rounds toward zero below 2^-24, the low half to nearest): the
// slab's A operands would be all zero and its MFMAs would add nothing.
// cb[2][kSlab] is the wave's coefficient double buffer (slab s in
// cb[s & 1]); cpre carries the prefetched pair.
__device__ __forceinline__ int skip_dead_slabs(double (*cb)[kSlab],
                                               const double *__restrict__ co, int lastc,
                                               int lane, double x, int m, int lmax, int nsl,
                                               double &p0, double &p1, int &k, double &cpre) {
  constexpr double kTiny = 0x1p-25;
  int s = 0;
  while (s < nsl && m + kSlab * (s + 1) - 1 <= lmax && wave_all(k < 0 || fabs(p1) < kTiny)) {
    const int cur = s & 1;
    cpre = co[min(kSlab * (s + 1) + (lane & (kSlab - 1)), lastc)];
    const double q0 = p0, q1 = p1;
    const int kq = k;
    double mx = 0.0;
#pragma unroll 8
    for (int j = 0; j < kSlab; ++j) {
      const double c = cb[cur][j];
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
    if (lane < kSlab) cb[cur ^ 1][lane] = cpre;
    __builtin_amdgcn_wave_barrier();
    if (!wave_all(kq < 0 || mx < kTiny)) {  // this slab has values: undo it
      p0 = q0;
      p1 = q1;
      k = kq;
      break;
    }
    ++s;
  }
  return s;
}

// The start table of a plan (ga_plan::d_f16_s0 / d_f16_st / d_f16_k): one
// wave per (m, group of 64 ring pairs), the synthesis kernel's preamble --
// the seed, the coefficient double buffer, skip_dead_slabs -- run once per
// plan instead of in every synthesis.  Lanes past the last ring pair repeat
// it, as in the synthesis.  The state is the one the synthesis loop would
// start from (the same operations in the same order), so it replaces the
// preamble there without changing a bit of the result.
__global__ __launch_bounds__(256) void legendre_f16_start_kernel(
    int lmax, int npairs, int ngroups, const double *__restrict__ z,
    const double *__restrict__ sth, const double *__restrict__ gam,
    const double *__restrict__ cm, int lam_exp, int32_t *__restrict__ s0tab,
    double2 *__restrict__ st, int32_t *__restrict__ stk) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int grp = int(blockIdx.x) * 4 + wave, m = int(blockIdx.y);
  if (grp >= ngroups) return;  // wave-uniform
  const int rp = grp * 64 + lane;
  const int rpc = rp < npairs ? rp : npairs - 1;
  const double x = z[rpc];
  double p0, p1 = 0.0;
  int k;
  legendre_seed(ldexp(cm[m], lam_exp), sth[rpc], m, p0, k);
  p0 = -p0;
  const double *__restrict__ co = gam + hp_index(lmax, m, m);
  __shared__ __align__(16) double cbuf[4][2][kSlab];
  const int lastc = lmax - m;
  double cpre = co[min(lane & (kSlab - 1), lastc)];
  if (lane < kSlab) cbuf[wave][0][lane] = cpre;
  __builtin_amdgcn_wave_barrier();
  const int nsl = int(nslab_of(lmax, m));
  const int s = skip_dead_slabs(cbuf[wave], co, lastc, lane, x, m, lmax, nsl, p0, p1, k, cpre);
  const int64_t i = int64_t(m) * (int64_t(ngroups) * 64) + rp;
  st[i] = make_double2(p0, p1);
  stk[i] = k;
  if (lane == 0) s0tab[int64_t(m) * ngroups + grp] = s;
}

// Two-chunk parking in the ring buffer.  The first chunk's raw accumulators
// of a wave are stored where the wave's own output goes: E of (row, col) in
// F at the north ring, O at the south ring (as store_tile addresses them;
// every (ring, m, column) of F belongs to exactly one wave), O of the
// equator row, whose north and south ring coincide, in eqo[m ncols + col].
// At the end each lane loads back what it stored (program order, same
// addresses) and adds it before the epilogue, which then overwrites F.  The
// result is the same as with the partial sums kept elsewhere; F itself is
// the parking space, so the chunked form needs no workspace beyond
// (lmax + 1) ncols floats.
template <int NT, bool LOAD>
__device__ __forceinline__ void park_tile(f32x16 (&acc)[2][2][NT], int m, int lmax, int npairs,
                                          int nrings, int ncols, int colbase, int rp0, int lane,
                                          float *__restrict__ F, float *__restrict__ eqo,
                                          int cshift) {
  const int half = lane >> 5, col_in = lane & 31;
  // Buffer addressing from two wave-uniform bases, so that a store costs no
  // address registers: north row rp0 + r at byte r rstride, south row
  // nrings - 1 - rp0 - r at byte (63 - r) rstride of its base; the column
  // part is per lane, the row part per element (scalar).  The host parks in
  // F only where 64 rows span less than 4 GiB.
  const uint32_t cg = 1u << cshift;  // real columns per group
  const uint32_t rstride = uint32_t((ncols + cg - 1) >> cshift) * uint32_t(lmax + 1) * cg * 4u;
  const int r0 = __builtin_amdgcn_readfirstlane(rp0);
  char *const fnb = reinterpret_cast<char *>(F) + int64_t(r0) * rstride + int64_t(m) * cg * 4;
  char *const fsb =
      reinterpret_cast<char *>(F) + (int64_t(nrings - 1 - r0) - 63) * rstride + int64_t(m) * cg * 4;
  const __amdgpu_buffer_rsrc_t rn = __builtin_amdgcn_make_buffer_rsrc(fnb, 0, -1, 0x00020000);
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(fsb, 0, -1, 0x00020000);
#pragma unroll
  for (int ct = 0; ct < NT; ++ct) {
    const int col = colbase + ct * 32 + col_in;
    const bool colok = col < ncols;
    const uint32_t cofs =
        uint32_t(col >> cshift) * uint32_t(lmax + 1) * cg * 4u + uint32_t(col & (cg - 1)) * 4u;
    const uint32_t vn = cofs + uint32_t(4 * half) * rstride;
    const uint32_t vs = cofs + uint32_t(4 * (1 - half)) * rstride;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int rr = 32 * t + 8 * (e >> 2) + (e & 3);  // row without the 4 half part
        const int rpo = r0 + rr + 4 * half;
        const int sn = rr * int(rstride), ss = (59 - rr) * int(rstride);
        if (colok && rpo < npairs) {
          if (LOAD) {
            acc[0][t][ct][e] += __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rn, vn, sn, 0));
            acc[1][t][ct][e] +=
                rpo < npairs - 1
                    ? __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, vs, ss, 0))
                    : eqo[int64_t(m) * ncols + col];
          } else {
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(acc[0][t][ct][e]), rn, vn, sn, 0);
            if (rpo < npairs - 1)
              __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(acc[1][t][ct][e]), rs, vs, ss,
                                                    0);
            else
              eqo[int64_t(m) * ncols + col] = acc[1][t][ct][e];
          }
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// The parking slot of a wave (in float4): one per (m, tile of the m, wave),
// numbered m-major.  Only m whose l-range reaches the split threshold park,
// and those are m < lmax + 1 - 32 (chunk_min - 1): the region holds that many
// m, not all of them.
__device__ __forceinline__ int64_t park_slot(int m, int rem, int G, int wave, int nt) {
  return ((int64_t(m) * G + rem) * 4 + wave) * (16 * nt * 64);
}

// Two-chunk parking in a region of its own: the wave's accumulators as 16 NT
// float4 per lane, lane-major (every store and load a whole 1 KiB per wave),
// in the wave's slot of the region (one slot per wave of the grid, 16 NT
// KiB).  Faster than parking in F (whose 32-byte column groups make every
// store and load a scatter), at the cost of a region as large as F.
template <int NT, bool LOAD>
__device__ __forceinline__ void park_lanes(f32x16 (&acc)[2][2][NT], float4 *__restrict__ slot) {
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) {
        float4 *const q4 = slot + ((p * 2 + t) * NT + ct) * 4 * 64;
        if (LOAD) {
          float4 o[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) o[q] = q4[q * 64];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            acc[p][t][ct][4 * q] += o[q].x;
            acc[p][t][ct][4 * q + 1] += o[q].y;
            acc[p][t][ct][4 * q + 2] += o[q].z;
            acc[p][t][ct][4 * q + 3] += o[q].w;
          }
          __builtin_amdgcn_sched_barrier(0);
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q)
            q4[q * 64] = make_float4(acc[p][t][ct][4 * q], acc[p][t][ct][4 * q + 1],
                                     acc[p][t][ct][4 * q + 2], acc[p][t][ct][4 * q + 3]);
        }
      }
}

// One geometry of the split-FP16 contraction: plan tables, the prepared B
// operands and column exponents, the ring buffer F and the parking space of
// the equator row's O half (two-chunk form).
struct f16_geom {
  const double *z, *sth, *gam, *cm;
  const int32_t *ring_pnum, *ring_pden;
  const int64_t *slab_off;
  const uint4 *B;
  const int32_t *colexp;
  float *F;
  float *eqo;     // two chunks parked in F: the equator row's O half
  float4 *part;   // two chunks parked in a region of their own (sep)
  const int32_t *s0tab;  // the plan's start table (ga_plan::d_f16_*)
  const double2 *st;
  const int32_t *stk;
  int ngroups;           // groups of 64 ring pairs
  int cshift;            // ring buffer: 1 << cshift real columns per group
  int sep;
  int lmax, npairs, nrings, ncols, nct, nrg, ncg, lam_exp, flush, nchunk, nt, blk_begin;
  int mgroup;     // consecutive m per XCD turn (1 or 4)
  int chunk_min;  // two-chunk form: least l-range (slabs) that is split
};

struct f16_group {
  f16_geom g[kGroupedMax];
  int ngroup;
};

// the value of the partner lane of a (re, im) pair: lanes 2c <-> 2c + 1
__device__ __forceinline__ float pair_swap(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));
}

// Epilogue of a wave's accumulators (as in the FP32 stage): D layout of
// 32x32: lane L, element e -> row 8 (e/4) + 4 (L/32) + e%4, column L % 32.
// Unscale by 2^-(8 + shift of the column), north = E + O, south = E - O,
// times the ring phase e^{i m phi0} (each lane computes its own ring's phase
// in FP64; rows fetch it by shuffle; the (re, im) columns of a shell sit in
// lanes (2c, 2c + 1), exchanged by DPP).  Ring buffer layout (cg = 1 << cshift
// real columns per group, ring_group): real column c of ring r at
//   ((r ngrp + c/cg)(lmax+1) + m) cg + c%cg,  ngrp = ceil(ncols / cg).
// Addresses: per element a wave-uniform ring base (the row's ring for the
// lower half-wave) plus a per-lane 32-bit offset fixed for the whole tile
// (the column part, and 4 rings for the other half-wave), so that a store
// costs no address arithmetic.
template <int NT>
__device__ __forceinline__ void store_tile(const f32x16 (&acc)[2][2][NT], int m, int lmax,
                                           int npairs, int nrings, int ncols, int colbase, int rp0,
                                           int rpc, int lane, const int32_t *__restrict__ ring_pnum,
                                           const int32_t *__restrict__ ring_pden,
                                           const int32_t *__restrict__ colexp,
                                           float *__restrict__ F, int lam_exp, int cshift) {
  const int half = lane >> 5, col_in = lane & 31;
  float phc, phs;
  {
    const int pnum = ring_pnum[rpc], pden = ring_pden[rpc];
    double ps = 0.0, pc = 1.0;
    if (pnum) sincospi(2.0 * double((int64_t(m) * pnum) % pden) / double(pden), &ps, &pc);
    phc = float(pc);
    phs = float(ps);
  }
  // rotation by (c, s): re' = re c - im s, im' = im c + re s, i.e. own value
  // times c plus the partner's times -s (real lane) or +s (imaginary lane);
  // at m = 0 the imaginary part is 0 (both factors 0 in an imaginary lane)
  const bool is_im = (col_in & 1) != 0;
  const float zf = (m == 0 && is_im) ? 0.0f : 1.0f;
  const float sg = is_im ? zf : -zf;
  const int cg = 1 << cshift;  // real columns per group
  const int64_t rstride = int64_t((ncols + cg - 1) >> cshift) * (lmax + 1) * cg;  // floats per ring
  float fac[NT];
  bool cok[NT];
  uint32_t noff[NT], soff[NT];
#pragma unroll
  for (int ct = 0; ct < NT; ++ct) {
    const int col = colbase + ct * 32 + col_in;
    cok[ct] = col < ncols;
    fac[ct] = cok[ct] ? ldexpf(1.0f, -(lam_exp + colexp[col])) : 0.0f;
    // byte offsets (< 2^32: the column part of a ring plus 4 rings)
    const uint32_t cofs = uint32_t((((col >> cshift) * (lmax + 1) + m) << cshift) + (col & (cg - 1))) * 4u;
    noff[ct] = cofs + uint32_t(half * 4 * rstride) * 4u;
    soff[ct] = cofs + uint32_t((1 - half) * 4 * rstride) * 4u;
  }
  const int rpu = __builtin_amdgcn_readfirstlane(rp0);  // wave-uniform: scalar ring bases
  // every row has its north and south ring and every lane its columns: the
  // stores need no guard
  const bool full = rpu + 63 < npairs - 1 && wave_all(cok[0] && cok[NT - 1]);
  auto tile = [&](auto guarded) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int r0 = 32 * t + 8 * (e >> 2) + (e & 3);  // row of the lower half-wave
        const int row = r0 + 4 * half;
        const float cz = __shfl(phc, row, 64) * zf, sz = __shfl(phs, row, 64) * sg;
        // north ring rpu + r0 (+ 4 in the upper half-wave); south ring
        // nrings - 1 - rpu - r0 - 4 (+ 4 in the lower half-wave)
        char *const fn = reinterpret_cast<char *>(F + int64_t(rpu + r0) * rstride);
        char *const fs = reinterpret_cast<char *>(F + int64_t(nrings - 5 - rpu - r0) * rstride);
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) {
          const float ev = acc[0][t][ct][e] * fac[ct], ov = acc[1][t][ct][e] * fac[ct];
          const float nv = ev + ov, sv = ev - ov;
          const float nrot = fmaf(pair_swap(nv), sz, nv * cz);
          const float srot = fmaf(pair_swap(sv), sz, sv * cz);
          if (!decltype(guarded)::value) {
            *reinterpret_cast<float *>(fn + noff[ct]) = nrot;
            *reinterpret_cast<float *>(fs + soff[ct]) = srot;
          } else {
            const int rpo = rpu + row;
            if (cok[ct] && rpo < npairs) *reinterpret_cast<float *>(fn + noff[ct]) = nrot;
            if (cok[ct] && rpo < npairs - 1) *reinterpret_cast<float *>(fs + soff[ct]) = srot;
          }
        }
      }
    }
  };
  if (full) tile(std::false_type{});
  else tile(std::true_type{});
}

// FLUSH: every `flush` slabs the accumulators are added into FP32 master sums
// and cleared, so that no accumulator chain is longer than `flush` slabs
// (3 flush MFMA roundings) and the master sums take one rounding per flush
// instead of 3 flush: the rounding error of the l-sums shrinks by about
// sqrt(1 / (3 flush) + flush / nslab), which the long sums of lmax > 4096
// need to stay within 1e-5 of the map RMS.  It doubles the accumulator
// registers, so it is instantiated with one column tile per wave.
//
// The body works on one geometry, described by an f16_geom, for block b of
// that geometry's grid; the plain kernel runs one geometry, the grouped
// kernel below several in one grid.
// SEP: the two-chunk form parks in a region of its own (park_lanes) instead
// of in F (park_tile); one variant per kernel keeps the register allocation
// of each within the 256 registers of two waves per SIMD.
template <int NT, bool FLUSH, bool SEP>
__device__ __forceinline__ void contract_blocks(const f16_geom &g, const int b) {
  const int lmax = g.lmax, npairs = g.npairs, nrings = g.nrings, ncols = g.ncols, nct = g.nct,
            nrg = g.nrg, ncg = g.ncg, lam_exp = g.lam_exp, flush = g.flush, nchunk = g.nchunk;
  const double *__restrict__ z = g.z;
  const double *__restrict__ gam = g.gam;
  const int32_t *__restrict__ ring_pnum = g.ring_pnum;
  const int32_t *__restrict__ ring_pden = g.ring_pden;
  const int64_t *__restrict__ slab_off = g.slab_off;
  const uint4 *__restrict__ B = g.B;
  const int32_t *__restrict__ colexp = g.colexp;
  float *__restrict__ F = g.F;
  float *__restrict__ eqo = g.eqo;
  // XCD-aware decode: block b runs on XCD b % 8 (a geometry's blocks start at
  // a multiple of 8 of the grid); XCD x takes m = x, x + 8, ... (largest
  // l-range first), all (ring group, column group) tiles of one m back to
  // back so they share that m's B slabs in the XCD's L2.
  const int xcd = b & 7, kq = b >> 3, G = nrg * ncg;
  const int mq = kq / G;  // this XCD's m slot
  const int m = g.mgroup == 4 ? 32 * (mq >> 2) + 4 * xcd + (mq & 3) : 8 * mq + xcd;
  if (m > lmax) return;
  const int rem = kq - mq * G;
  const int rg = rem / ncg, cg = rem - (rem / ncg) * ncg;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int rp0 = rg * 256 + wave * 64;
  if (rp0 >= npairs) return;  // wave-uniform; the kernel has no barriers
  const int rp = rp0 + lane;
  const int rpc = rp < npairs ? rp : npairs - 1;
  const double x = z[rpc];
  // the normalised recurrence y_l = (gam_l x) y_{l-1} - y_{l-2}: one multiply
  // and one FMA per step, the FMA alone on the dependency chain.  It starts
  // at l = m from (-seed, 0) (gam at l = m is 0, so one step yields
  // y_m = lambda_mm); the leading slabs in which every lane stays below the
  // FP16 range of the split (near the poles at large m) contribute nothing,
  // and the plan's start table holds, per wave, the first slab that does not
  // (skip_dead_slabs, run at plan creation) and each lane's state there.
  double p0, p1;  // y_{l-2}, y_{l-1} (times 2^lam_exp)
  int k;          // scale count (value = p 2^(600 k))
  int s;          // first slab run in full
  {
    const int64_t si = int64_t(m) * (int64_t(g.ngroups) * 64) + rp;
    const double2 st = g.st[si];
    p0 = st.x;
    p1 = st.y;
    k = g.stk[si];
    s = __builtin_amdgcn_readfirstlane(g.s0tab[int64_t(m) * g.ngroups + (rp0 >> 6)]);
  }

  const int64_t base = hp_index(lmax, m, m);
  const double *__restrict__ co = gam + base;
  const int64_t s0 = slab_off[m];
  const int nsl = int(nslab_of(lmax, m));
  const int half = lane >> 5, col_in = lane & 31;
  const int colbase = cg * 32 * NT;
  // column tiles past the last real one re-read the last tile (never stored)
  int ctile[NT];
#pragma unroll
  for (int ct = 0; ct < NT; ++ct) ctile[ct] = min(cg * NT + ct, nct - 1);

  f32x16 acc[2][2][NT];  // [parity][row tile][col tile]
  f32x16 mst[2][2][FLUSH ? NT : 1];  // master sums of the flushing form
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) {
        acc[p][t][ct] = f32x16{};
        if (FLUSH) mst[p][t][FLUSH ? ct : 0] = f32x16{};
      }
  int since_flush = 0;  // slabs accumulated since the last flush

  // Recurrence coefficients go through LDS, one slab (32 double2) per wave
  // and buffer: the next slab's are loaded from global memory (lanes 0..31,
  // one pair each) while the current slab runs, then written to the other
  // buffer; every step reads its pair with a broadcast LDS load.  Only this
  // wave touches its buffers, so no workgroup barrier is needed.
  __shared__ __align__(16) double cbuf[4][2][kSlab];
  // Coefficient and B-operand loads: a wave-uniform base plus a 32-bit lane
  // offset fixed for the whole loop (no per-slab address arithmetic).  The
  // coefficients are not clamped at the end of m: the plan pads gam, and
  // rows past this m's last one only feed steps whose values are discarded.
  // (buffer loads: scalar base and offset per slab, the lane's offset in a
  // VGPR; this m's B operands span less than 4 GiB)
  const uint32_t coff = uint32_t(lane & (kSlab - 1)) * 8u, boff = uint32_t(lane) * 16u;
  const __amdgpu_buffer_rsrc_t crs =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(co), 0, -1, 0x00020000);
  const __amdgpu_buffer_rsrc_t brs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<uint4 *>(B + s0 * 2 * nct * 128), 0, -1, 0x00020000);
  double cpre = __builtin_bit_cast(
      double, __builtin_amdgcn_raw_buffer_load_b64(crs, coff, kSlab * 8 * s, 0));
  if (lane < kSlab) cbuf[wave][s & 1][lane] = cpre;
  __builtin_amdgcn_wave_barrier();

  // the slab loop from s to end (accumulating into acc)
  auto run_slabs = [&](const int end) {
    for (; s < end; ++s) {
      const int l0 = m + kSlab * s;
      const int cur = s & 1;
      // next slab's coefficients (past the end of m: never used), issued
      // first: they are written to LDS before the MFMAs need the B operands
      cpre = __builtin_bit_cast(
          double, __builtin_amdgcn_raw_buffer_load_b64(crs, coff, kSlab * 8 * (s + 1), 0));
      // B operands of this slab, consumed after the recurrence below
      uint4 bh[2][NT], bl[2][NT];
  #pragma unroll
      for (int p = 0; p < 2; ++p)
  #pragma unroll
        for (int ct = 0; ct < NT; ++ct) {
          const int so = ((s * 2 + p) * nct + ctile[ct]) * 2048;
          bh[p][ct] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(brs, boff, so, 0));
          bl[p][ct] =
              __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(brs, boff + 1024, so, 0));
        }
      const double *__restrict__ cb = cbuf[wave][cur];

      uint32_t ph[2][8], pl[2][8];  // [parity][group]: packed (k = 2g, 2g + 1)
      slab_values(cb, x, l0, lmax, p0, p1, k, ph, pl);
      if (lane < kSlab) cbuf[wave][cur ^ 1][lane] = cpre;
      __builtin_amdgcn_wave_barrier();

      // MFMAs only if some lane has a nonzero lambda in this slab (k only
      // grows, so a lane still scaled at the end of the slab contributed nothing)
      if (wave_any(k == 0)) {
        f16::slab_mfma<NT>(ph, pl, bh, bl, acc);
        if (FLUSH && ++since_flush == flush) {
  #pragma unroll
          for (int p = 0; p < 2; ++p)
  #pragma unroll
            for (int t = 0; t < 2; ++t)
  #pragma unroll
              for (int ct = 0; ct < NT; ++ct) {
                mst[p][t][FLUSH ? ct : 0] += acc[p][t][ct];
                acc[p][t][ct] = f32x16{};
              }
          since_flush = 0;
        }
      }
    }
  };

  // Two-chunk form (nchunk == 2): the l-range left after the dead slabs, if
  // at least g.chunk_min slabs long, is cut in two halves; the first half's accumulators are parked in F
  // (park_tile), the second half starts from zero, and the parked sums are
  // added back before the epilogue: two shorter FP32 rounding chains.  The
  // two halves are two copies of the slab loop, not one loop with a cut in
  // it, so the parking code sits outside every loop.
  const bool two = !FLUSH && nchunk == 2 && nsl - s >= g.chunk_min;
  run_slabs(two ? s + (nsl - s) / 2 : nsl);
  if (two) {
    if (SEP) park_lanes<NT, false>(acc, g.part + park_slot(m, rem, G, wave, NT) + lane);
    else park_tile<NT, false>(acc, m, lmax, npairs, nrings, ncols, colbase, rp0, lane, F, eqo, g.cshift);
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) acc[p][t][ct] = f32x16{};
    run_slabs(nsl);
    if (SEP) park_lanes<NT, true>(acc, g.part + park_slot(m, rem, G, wave, NT) + lane);
    else park_tile<NT, true>(acc, m, lmax, npairs, nrings, ncols, colbase, rp0, lane, F, eqo, g.cshift);
  }
  if (FLUSH) {
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) acc[p][t][ct] += mst[p][t][FLUSH ? ct : 0];
  }
  store_tile<NT>(acc, m, lmax, npairs, nrings, ncols, colbase, rp0, rpc, lane, ring_pnum, ring_pden,
                 colexp, F, lam_exp, g.cshift);
}

template <int NT, int MINW, bool FLUSH, bool SEP>
__global__ __launch_bounds__(256, MINW) void legendre_synth_f16_kernel(const f16_geom g) {
  contract_blocks<NT, FLUSH, SEP>(g, int(blockIdx.x));
}

// Several geometries in one grid (an uneven batch): geometry i owns blocks
// [blk_begin_i, blk_begin_{i+1}), largest first, so its long blocks are
// dispatched first and the short blocks of the small geometries fill the
// CUs as the large ones drain.  The geometry of a block is uniform, and so is
// its tile count (one or two 32-column tiles per wave, both bodies are
// compiled in).  One accumulation chain or l-chunks, no register master sums.
template <int MINW, bool SEP>
__global__ __launch_bounds__(256, MINW) void legendre_synth_f16_grouped_kernel(
    const f16_group grp) {
  const int bx = int(blockIdx.x);
  int gi = 0;
#pragma unroll
  for (int i = 1; i < kGroupedMax; ++i)
    if (i < grp.ngroup && bx >= grp.g[i].blk_begin) gi = i;
  f16_geom g = grp.g[0];
#pragma unroll
  for (int i = 1; i < kGroupedMax; ++i)
    if (i == gi) g = grp.g[i];
  const int b = bx - g.blk_begin;
  if (g.nt == 2) contract_blocks<2, false, SEP>(g, b);
  else contract_blocks<1, false, SEP>(g, b);
}

// Producer / consumer form.  In the single-wave form every wave runs the
// FP64 recurrence (VALU) and then its MFMAs; the two waves of a SIMD do the
// same work in step, so the two phases add up instead of overlapping.  Here
// a workgroup of 8 waves splits the roles on 4 groups of 64 ring pairs:
// wave w (producer) runs the recurrence and the FP16 splitting and passes
// each slab's A operands through LDS (double-buffered, one workgroup barrier
// per slab); wave w + 4 (consumer) runs the MFMAs of that slab against its
// 64 columns while the producer computes the next slab, and stores the
// result.  The consumer keeps the B operands of the next two slabs in
// registers, so their loads have two slab periods to arrive.
__global__ __launch_bounds__(512, 1) void legendre_synth_f16_pc_kernel(
    int lmax, int npairs, int nrings, int ncols, int nct, int nrg, int ncg,
    const double *__restrict__ z, const double *__restrict__ sth, const double *__restrict__ gam,
    const double *__restrict__ cm, const int32_t *__restrict__ ring_pnum,
    const int32_t *__restrict__ ring_pden, const int64_t *__restrict__ slab_off,
    const uint4 *__restrict__ B, const int32_t *__restrict__ colexp, float *__restrict__ F,
    int lam_exp) {
  constexpr int NT = 2;
  // XCD-aware decode as in the single-wave form (ncg counts 64-column groups)
  const int b = blockIdx.x;
  const int xcd = b & 7, kq = b >> 3, G = nrg * ncg;
  const int m = 8 * (kq / G) + xcd;
  if (m > lmax) return;  // whole workgroup
  const int rem = kq - (kq / G) * G;
  const int rg = rem / ncg, cg = rem - (rem / ncg) * ncg;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int pr = wave & 3;  // ring-pair group of the pair
  const bool producer = wave < 4;
  const int rp0 = rg * 256 + pr * 64;
  const bool active = rp0 < npairs;  // same for both waves of a pair
  const int rp = rp0 + lane;
  const int rpc = rp < npairs ? rp : npairs - 1;
  const int64_t base = hp_index(lmax, m, m);
  const int64_t s0 = slab_off[m];
  const int nsl = int(nslab_of(lmax, m));

  // A operands of a slab, [buffer][pair][hi p0t0 p0t1 p1t0 p1t1 | lo ...][lane],
  // and whether the slab has any nonzero lambda
  __shared__ __align__(16) uint4 abuf[2][4][8][64];
  __shared__ int aflag[2][4];

  if (producer) {
    __shared__ __align__(16) double cbuf[4][2][kSlab];
    const double x = z[rpc];
    double p0, p1 = 0.0;
    int k;
    legendre_seed(ldexp(cm[m], lam_exp), sth[rpc], m, p0, k);
    p0 = -p0;
    const double *__restrict__ co = gam + base;
    const int lastc = lmax - m;
    double cpre = co[min(lane & (kSlab - 1), lastc)];
    if (lane < kSlab) cbuf[pr][0][lane] = cpre;
    __builtin_amdgcn_wave_barrier();
    for (int s = 0; s < nsl; ++s) {
      const int cur = s & 1;
      if (active) {
        cpre = co[min(kSlab * (s + 1) + (lane & (kSlab - 1)), lastc)];
        uint32_t ph[2][8], pl[2][8];
        slab_values(cbuf[pr][cur], x, m + kSlab * s, lmax, p0, p1, k, ph, pl);
        if (lane < kSlab) cbuf[pr][cur ^ 1][lane] = cpre;
        __builtin_amdgcn_wave_barrier();
        const bool need = wave_any(k == 0);
        if (need) {
          uint4 ah[2][2], al[2][2];
          f16::slab_a(ph, pl, ah, al);
#pragma unroll
          for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int t = 0; t < 2; ++t) {
              abuf[cur][pr][2 * p + t][lane] = ah[p][t];
              abuf[cur][pr][4 + 2 * p + t][lane] = al[p][t];
            }
        }
        if (lane == 0) aflag[cur][pr] = need ? 1 : 0;
      }
      __syncthreads();  // slab s's A operands are in LDS
    }
    return;
  }

  // consumer
  const int colbase = cg * 32 * NT;
  int ctile[NT];
#pragma unroll
  for (int ct = 0; ct < NT; ++ct) ctile[ct] = min(cg * NT + ct, nct - 1);
  f32x16 acc[2][2][NT];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) acc[p][t][ct] = f32x16{};
  // B of slabs s (bc) and s + 1 (bn); slab indices clamped at the end
  uint4 bch[2][NT], bcl[2][NT], bnh[2][NT], bnl[2][NT];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) {
      const uint4 *__restrict__ u0 = B + ((s0 * 2 + p) * nct + ctile[ct]) * 128;
      const uint4 *__restrict__ u1 = B + (((s0 + min(1, nsl - 1)) * 2 + p) * nct + ctile[ct]) * 128;
      bch[p][ct] = u0[lane];
      bcl[p][ct] = u0[64 + lane];
      bnh[p][ct] = u1[lane];
      bnl[p][ct] = u1[64 + lane];
    }
  for (int s = 0; s < nsl; ++s) {
    const int cur = s & 1;
    __syncthreads();  // the producer has written slab s
    if (active && aflag[cur][pr] != 0) {
      uint4 ah[2][2], al[2][2];
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          ah[p][t] = abuf[cur][pr][2 * p + t][lane];
          al[p][t] = abuf[cur][pr][4 + 2 * p + t][lane];
        }
      f16::slab_mfma_a<NT>(ah, al, bch, bcl, acc);
    }
    // rotate the B pipeline and request slab s + 2
    const int sn = min(s + 2, nsl - 1);
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) {
        bch[p][ct] = bnh[p][ct];
        bcl[p][ct] = bnl[p][ct];
        const uint4 *__restrict__ u = B + (((s0 + sn) * 2 + p) * nct + ctile[ct]) * 128;
        bnh[p][ct] = u[lane];
        bnl[p][ct] = u[64 + lane];
      }
  }
  if (!active) return;
  store_tile<NT>(acc, m, lmax, npairs, nrings, ncols, colbase, rp0, rpc, lane, ring_pnum, ring_pden,
                 colexp, F, lam_exp, ring_group_cshift(ncols / 2));
}

}  // namespace

int launch_legendre_prepare_f16(const ga_plan *p, int S, const float *alm, void *bws,
                                int32_t *colexp, hipStream_t stream, int lmin,
                                const float *colpart, int64_t nparts, const double *csc) {
  const int ncols = 2 * S;
  const int nct = (ncols + 31) / 32;
  uint4 *B = reinterpret_cast<uint4 *>(bws);
  // per-column maxima scratch, right after the nct * 32 column exponents
  uint32_t *cmax = reinterpret_cast<uint32_t *>(colexp) + nct * 32;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(unsigned((ncols + 255) / 256)), dim3(256), 0, stream,
                     cmax, ncols, 0u);
  GA_LAUNCH_CHECK();
  // the column maxima over the alm rows, or over per-block maxima of them
  // (same ncols; the maximum of maxima is the maximum)
  const float *src = colpart ? colpart : alm;
  const int64_t nrows = colpart ? nparts : p->nalm;
  int nblk = int(std::min<int64_t>((nrows + 255) / 256, 2048));
  if (ncols % 4 == 0 && ncols <= 1024 && (reinterpret_cast<uintptr_t>(src) & 15) == 0)
    hipLaunchKernelGGL(colmax4_kernel, dim3(unsigned(nblk)), dim3(256), 0, stream, nrows, ncols,
                       reinterpret_cast<const float4 *>(src), cmax);
  else
    hipLaunchKernelGGL(colmax_kernel, dim3(unsigned(nblk)), dim3(256), 0, stream, nrows, ncols,
                       src, cmax);
  GA_LAUNCH_CHECK();
  GA_REQUIRE(p->lmax < 65535, GA_ERR_UNSUPPORTED, "lmax too large for the operand preparation");
  const int units0 = int(nslab_of(p->lmax, 0)) * 2 * nct;
  hipLaunchKernelGGL(swizzle_b_kernel, dim3(unsigned((units0 + 3) / 4), unsigned(p->lmax + 1)),
                     dim3(256), 0, stream, p->lmax, lmin, ncols, nct, p->d_slab_off, alm, cmax,
                     csc, B, colexp);
  GA_LAUNCH_CHECK();
  return GA_OK;
}

// the f16_geom of plan p for S shells (flush resolved, nt 1 or 2) and its
// block count
// whether the epilogue's 32-bit per-lane byte offsets cover the ring buffer
// rows of S shells: a ring's columns plus 4 rings within 4 GiB
static bool f16_offsets_fit(const ga_plan *p, int S) {
  const uint64_t cg = uint64_t(2 * ring_group(S));
  const uint64_t ring_bytes = (uint64_t(2 * S) + cg - 1) / cg * cg * uint64_t(p->lmax + 1) * 4;
  return 5 * ring_bytes < (uint64_t(1) << 32);
}

static f16_geom f16_geometry(const ga_plan *p, int S, float *F, const void *bws,
                             const int32_t *colexp, int nt, int flush, void *park,
                             int mgroup, int chunk_min, int64_t *nblk) {
  f16_geom g{};
  g.z = p->d_z;
  g.sth = p->d_sth;
  g.gam = p->d_gam;
  g.cm = p->d_cm;
  g.s0tab = p->d_f16_s0;
  g.st = p->d_f16_st;
  g.stk = p->d_f16_k;
  g.ngroups = p->f16_ngroups;
  g.cshift = ring_group_cshift(S);
  g.ring_pnum = p->d_ring_pnum;
  g.ring_pden = p->d_ring_pden;
  g.slab_off = p->d_slab_off;
  g.B = reinterpret_cast<const uint4 *>(bws);
  g.colexp = colexp;
  g.F = F;
  g.sep = flush == -3;
  g.eqo = flush == -2 ? static_cast<float *>(park) : nullptr;
  g.part = flush == -3 ? static_cast<float4 *>(park) : nullptr;
  g.lmax = p->lmax;
  g.npairs = int(p->npairs);
  g.nrings = int(p->nrings);
  g.ncols = 2 * S;
  g.nct = (g.ncols + 31) / 32;
  g.nrg = int((p->npairs + 255) / 256);
  g.ncg = (g.ncols + 32 * nt - 1) / (32 * nt);
  g.lam_exp = p->lam_exp16;
  g.flush = flush;
  g.nchunk = (flush == -2 || flush == -3) ? 2 : 1;
  g.nt = nt;
  g.blk_begin = 0;
  g.mgroup = mgroup == 1 ? 1 : 4;  // 0: automatic = 4
  g.chunk_min = chunk_min > 0 ? chunk_min : kTwoChunkMinSlabs;
  // m slots per XCD: enough for m = 0 .. lmax (a multiple of 4 with runs of 4)
  const int slots = g.mgroup == 4 ? 4 * ((p->lmax + 1 + 31) / 32) : (p->lmax + 1 + 7) / 8;
  *nblk = int64_t(8) * slots * g.nrg * g.ncg;
  return g;
}

int launch_legendre_f16_start(const ga_plan *p, hipStream_t stream) {
  GA_REQUIRE(p->d_f16_s0 && p->d_f16_st && p->d_f16_k, GA_ERR_INVALID,
             "start table not allocated");
  const int ngroups = p->f16_ngroups;
  GA_REQUIRE(ngroups * 64 >= p->npairs, GA_ERR_INVALID, "start table groups %d for %d ring pairs",
             ngroups, p->npairs);
  hipLaunchKernelGGL(legendre_f16_start_kernel, dim3(unsigned((ngroups + 3) / 4), unsigned(p->lmax + 1)),
                     dim3(256), 0, stream, p->lmax, int(p->npairs), ngroups, p->d_z, p->d_sth,
                     p->d_gam, p->d_cm, p->lam_exp16, p->d_f16_s0, p->d_f16_st, p->d_f16_k);
  GA_LAUNCH_CHECK();
  return GA_OK;
}

int launch_legendre_contract_f16(const ga_plan *p, int S, float *F, const void *bws,
                                 const int32_t *colexp, hipStream_t stream, int nt, int flush,
                                 void *park, int mgroup, int chunk_min) {
  const int ncols = 2 * S;
  const int nct = (ncols + 31) / 32;
  const uint4 *B = reinterpret_cast<const uint4 *>(bws);
  // flush (resolved): n > 0 master sums in registers (they take the registers
  // of a second tile); -2 / -3: two l-chunks, the first parked in F / in a
  // region of its own
  GA_REQUIRE(flush >= -3 && flush != 0, GA_ERR_INVALID, "accumulation form %d out of range",
             flush);
  GA_REQUIRE(flush != -3 || park, GA_ERR_INVALID, "two-chunk parking needs its region");
  GA_REQUIRE(flush != -2 || (park && f16_park_fits(p, S)), GA_ERR_UNSUPPORTED,
             "two-chunk accumulation needs 64 ring rows of %d shells within 4 GiB; "
             "use one chain (-1) or register master sums (n > 0)", S);
  GA_REQUIRE(f16_offsets_fit(p, S), GA_ERR_UNSUPPORTED,
             "%d shells at lmax %d exceed the synthesis epilogue's 32-bit offsets", S, p->lmax);
  if (flush > 0) nt = 1;
  if (nt != 1 && nt != 2 && nt != 4) nt = ncols <= 32 ? 1 : 2;
  if (nt == 4) {
    // producer / consumer wave pairs over 64-column groups
    const int nrg = (p->npairs + 255) / 256;
    const int ncg = (ncols + 63) / 64;
    const int nm8 = (p->lmax + 1 + 7) / 8;
    const int64_t nblk = int64_t(8) * nm8 * nrg * ncg;
    GA_REQUIRE(nblk < (int64_t(1) << 31), GA_ERR_UNSUPPORTED, "grid too large");
    hipLaunchKernelGGL(legendre_synth_f16_pc_kernel, dim3(unsigned(nblk)), dim3(512), 0, stream,
                       p->lmax, p->npairs, p->nrings, ncols, nct, nrg, ncg, p->d_z, p->d_sth,
                       p->d_gam, p->d_cm, p->d_ring_pnum, p->d_ring_pden, p->d_slab_off, B, colexp,
                       F, p->lam_exp16);
    GA_LAUNCH_CHECK();
    return GA_OK;
  }
  int64_t nblk2 = 0;
  const f16_geom g =
      f16_geometry(p, S, F, bws, colexp, nt, flush, park, mgroup, chunk_min, &nblk2);
  GA_REQUIRE(nblk2 < (int64_t(1) << 31), GA_ERR_UNSUPPORTED, "grid too large");
#define GA_F16_LAUNCH(NT_, MW_, FL_, SEP_)                                                      \
  hipLaunchKernelGGL((legendre_synth_f16_kernel<NT_, MW_, FL_, SEP_>), dim3(unsigned(nblk2)),  \
                     dim3(256), 0, stream, g)
  if (flush > 0) GA_F16_LAUNCH(1, 2, true, false);
  else if (nt == 1 && flush == -3) GA_F16_LAUNCH(1, 2, false, true);
  else if (nt == 1) GA_F16_LAUNCH(1, 2, false, false);
  else if (flush == -3) GA_F16_LAUNCH(2, 2, false, true);
  else GA_F16_LAUNCH(2, 2, false, false);
#undef GA_F16_LAUNCH
  GA_LAUNCH_CHECK();
  return GA_OK;
}

int launch_legendre_contract_grouped(int ngroup, const ga_plan *const *plans, const int32_t *S,
                                     void *const *ws, hipStream_t stream) {
  GA_REQUIRE(ngroup >= 1 && ngroup <= kGroupedMax, GA_ERR_INVALID,
             "a grouped launch takes 1 to %d geometries", kGroupedMax);
  f16_group grp{};
  grp.ngroup = ngroup;
  int64_t begin = 0;
  for (int i = 0; i < ngroup; ++i) {
    const ga_plan *p = plans[i];
    const ga_tuning tu = p->tune;
    GA_REQUIRE(tu.mfma_path != 1, GA_ERR_UNSUPPORTED,
               "geometry %d is tuned to the FP32 MFMA path; a grouped launch runs split FP16", i);
    const int acc_form = f16_accumulation(p, tu, S[i]);
    GA_REQUIRE(acc_form < 0, GA_ERR_UNSUPPORTED,
               "geometry %d accumulates in register master sums (lmax %d); a grouped launch "
               "runs one chain or l-chunks",
               i, p->lmax);
    const f32_ws_layout w = f32_workspace(p, S[i], tu);
    GA_REQUIRE(acc_form != -2 || (w.p_bytes > 0 && f16_park_fits(p, S[i])), GA_ERR_UNSUPPORTED,
               "geometry %d: two-chunk accumulation does not fit %d shells", i, S[i]);
    char *base = static_cast<char *>(ws[i]);
    int nt = tu.mfma_nt;
    if (nt != 1 && nt != 2) nt = 2 * S[i] <= 32 ? 1 : 2;
    int64_t nblk = 0;
    GA_REQUIRE(f16_offsets_fit(p, S[i]), GA_ERR_UNSUPPORTED,
               "geometry %d: %d shells at lmax %d exceed the epilogue's 32-bit offsets", i, S[i],
               p->lmax);
    grp.g[i] = f16_geometry(p, S[i], reinterpret_cast<float *>(base), base + w.b_off,
                            reinterpret_cast<const int32_t *>(base + w.e_off), nt, acc_form,
                            base + w.p_off, tu.m_group, tu.chunk_min, &nblk);
    grp.g[i].blk_begin = int(begin);  // a multiple of 8: XCD mapping as in one launch
    begin += nblk;
    GA_REQUIRE(begin < (int64_t(1) << 31), GA_ERR_UNSUPPORTED, "grid too large");
  }
  // one parking variant per grid: the geometries' two-chunk forms must agree
  bool sep = false, inf = false;
  for (int i = 0; i < ngroup; ++i) {
    if (grp.g[i].nchunk != 2) continue;
    (grp.g[i].sep ? sep : inf) = true;
  }
  GA_REQUIRE(!(sep && inf), GA_ERR_UNSUPPORTED,
             "a grouped launch takes one two-chunk parking form (-2 or -3), not both");
  if (sep)
    hipLaunchKernelGGL((legendre_synth_f16_grouped_kernel<2, true>), dim3(unsigned(begin)),
                       dim3(256), 0, stream, grp);
  else
    hipLaunchKernelGGL((legendre_synth_f16_grouped_kernel<2, false>), dim3(unsigned(begin)),
                       dim3(256), 0, stream, grp);
  GA_LAUNCH_CHECK();
  return GA_OK;
}

int launch_legendre_prepare_batch(const ga_plan *p, const ga_tuning &tu, int S, const float *alm,
                                  void *ws, hipStream_t stream, const float *colpart,
                                  int64_t nparts) {
  if (tu.mfma_path == 1) return GA_OK;  // the FP32 path reads the alm directly
  const f32_ws_layout w = f32_workspace(p, S, tu);
  char *base = static_cast<char *>(ws);
  return launch_legendre_prepare_f16(p, S, alm, base + w.b_off,
                                     reinterpret_cast<int32_t *>(base + w.e_off), stream, 0,
                                     colpart, nparts, p->d_csc);
}

int launch_legendre_contract_batch(const ga_plan *p, const ga_tuning &tu, int S,
                                   const float *alm, void *ws, hipStream_t stream) {
  float *F = static_cast<float *>(ws);
  if (tu.mfma_path == 1) return launch_legendre_synth_mfma(p, tu, S, alm, F, stream);
  const f32_ws_layout w = f32_workspace(p, S, tu);
  const char *base = static_cast<const char *>(ws);
  const int acc_form = f16_accumulation(p, tu, S);
  return launch_legendre_contract_f16(
      p, S, F, base + w.b_off, reinterpret_cast<const int32_t *>(base + w.e_off), stream,
      tu.mfma_nt, acc_form, const_cast<char *>(base) + w.p_off, tu.m_group, tu.chunk_min);
}

int launch_legendre_synth_batch(const ga_plan *p, int S, const float *alm, void *ws,
                                hipStream_t stream) {
  return launch_legendre_synth_batch(p, p->tune, S, alm, ws, stream);  // one snapshot
}

int launch_legendre_synth_batch(const ga_plan *p, const ga_tuning tu, int S, const float *alm,
                                void *ws, hipStream_t stream) {
  int rc = launch_legendre_prepare_batch(p, tu, S, alm, ws, stream);
  if (rc) return rc;
  return launch_legendre_contract_batch(p, tu, S, alm, ws, stream);
}

}  // namespace ga
