// Shared host/device helpers for the glass_amd HIP library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "glass_amd.h"

namespace ga {

// ---- error reporting ------------------------------------------------------
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();

#define GA_CHECK_HIP(expr)                                                     \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      ::ga::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),   \
                      __FILE__, __LINE__);                                     \
      return GA_ERR_HIP;                                                       \
    }                                                                          \
  } while (0)

#define GA_REQUIRE(cond, code, ...)                                            \
  do {                                                                         \
    if (!(cond)) {                                                             \
      ::ga::set_error(__VA_ARGS__);                                            \
      return (code);                                                           \
    }                                                                          \
  } while (0)

#define GA_LAUNCH_CHECK()                                                      \
  do {                                                                         \
    hipError_t e_ = hipGetLastError();                                         \
    if (e_ != hipSuccess) {                                                    \
      ::ga::set_error("kernel launch failed: %s (%s:%d)",                      \
                      hipGetErrorString(e_), __FILE__, __LINE__);              \
      return GA_ERR_HIP;                                                       \
    }                                                                          \
  } while (0)

// ---- tuning knobs: a process default (ga_set_tuning) copied into each plan
// at creation, then per plan (ga_plan_set_tuning); launchers read the plan's
// set, once per entry-point call ----
struct ga_tuning {
  int mfma_nt = 0;      // 32-column tiles per wave in the MFMA Legendre stage (0 = auto; 4 = wave pairs)
  int mfma_chunk = 16;  // quads per MFMA accumulation chunk (0 = single accumulator)
  int mfma_path = 0;    // Legendre stage: 0 auto (= split FP16), 1 FP32 MFMA, 2 split FP16
  int ring_path = 0;    // FP32 ring synthesis: 0 auto (fused passes, register fold on wide rings), 1 radix-4 stages, 2 fused passes only
  int spin_path = 0;    // FP32 spin synthesis: 0 auto (= split FP16 MFMA), 1 VALU, 2 split FP16 MFMA
  int anal_path = 0;    // FP32 analysis Legendre stage: 0 auto (= split FP16 MFMA), 1 VALU, 2 MFMA
  int m_group = 0;      // split-FP16 synthesis: consecutive m per XCD turn, 0 auto (= 4), 1, 4
  int chunk_min = 0;    // two-chunk form: split only l-ranges of at least this many slabs
                        // (0 auto = kTwoChunkMinSlabs)
  int f16_flush = 0;    // split-FP16 synthesis accumulation: n > 0 master sums in registers
                        // flushed every n slabs (one column tile), -2 / -3 two l-chunks
                        // (the first parked in the ring buffer / in a region of its own,
                        // faster, as large as the ring buffer), -1 one chain, 0 auto
                        // (kAutoFlushSlabs when lmax > kAutoFlushLmax, else kAutoChunks)
};
// lmax above which the split-FP16 synthesis flushes its MFMA accumulators into
// register master sums by default, the flush period in slabs, and the default
// accumulation below that lmax
constexpr int kAutoFlushLmax = 4096;
constexpr int kAutoFlushSlabs = 8;
constexpr int kAutoChunks = -3;
// The two-chunk form splits the l-range of a wave only when at least this
// many slabs (of 32 rows) remain after the dead ones: the long chains of the
// small m set the FP32 rounding error of a map, shorter ones add little to
// it.  Measured at nside 2048, 64 shells: split from 8 slabs 33.3 ms, worst
// shell 6.8e-6; from 32: 33.0 ms, 6.7e-6; from 64: 32.4 ms, 7.2e-6; from 96:
// 31.9 ms, 7.2e-6 (median 5.8e-6 against 5.3e-6).
constexpr int kTwoChunkMinSlabs = 64;
ga_tuning &tuning();  // the process default

// The FP32 ring buffer of a batch of S shells holds them in groups of
// G = ring_group(S): complex element (ring, m, s) at
//   ((ring ceil(S/G) + s/G)(lmax+1) + m) G + s%G.
// G = 8 for batches of at least 8 shells: each (ring, group, m) row is then
// 16 real columns, 64 bytes, so the synthesis epilogue's store of a
// half-wave's 32 columns touches two rows (G = 4: four 32-byte rows, 25 % slower
// contraction; G = 16: one 128-byte row, but the ring stage, whose workgroups
// read one shell pair across all m, touches a line per m and is slower by
// more); smaller batches use G = 4.  Measured at nside 2048, 64 shells:
// contraction + ring stage 42.4 (G = 4), 38.6 (G = 8), 40.2 ms (G = 16).
__host__ __device__ inline int ring_group(int nshell) { return nshell >= 8 ? 8 : 4; }
// log2 of the real columns (2 G) per group
__host__ __device__ inline int ring_group_cshift(int nshell) { return nshell >= 8 ? 4 : 3; }

// l rows per slab of the FP16-split MFMA Legendre stage (K = 16 per parity)
constexpr int kSlab = 32;
__host__ __device__ inline int64_t nslab_of(int64_t lmax, int64_t m) { return (lmax - m + kSlab) / kSlab; }

// Workspace of the FP32 batched synthesis (sizes for nbatch maps, ncols = 2
// nbatch real columns): the ring buffer F, then the split-FP16 B operands of
// the Legendre stage, then one int32 exponent per column, then (when the
// plan's synthesis accumulates in l-chunks) the partial sums of every wave.
// Each part starts at a multiple of 256 bytes.
struct f32_ws_layout {
  size_t f_bytes, b_off, b_bytes, e_off, p_off, p_bytes, total;
};
f32_ws_layout f32_workspace(const ga_plan *p, int nbatch, const ga_tuning &tu);
// under the plan's current tuning
f32_ws_layout f32_workspace(const ga_plan *p, int nbatch);
// the accumulation form the split-FP16 synthesis of plan p runs under tuning
// set tu: the f16_flush knob with 0 (automatic) resolved
int f16_accumulation(const ga_plan *p, const ga_tuning &tu, int nbatch);
// whether the two-chunk form can park in the ring buffer of nbatch shells
// (64 rows of it must span less than 4 GiB: 32-bit buffer offsets)
bool f16_park_fits(const ga_plan *p, int nbatch);

// ---- sizes ----------------------------------------------------------------
inline int64_t nalm_of(int64_t lmax) { return (lmax + 1) * (lmax + 2) / 2; }
inline int64_t npix_of(int64_t nside) { return 12 * nside * nside; }
inline int64_t nrings_of(int64_t nside) { return 4 * nside - 1; }
__host__ __device__ inline int64_t hp_index(int64_t lmax, int64_t l, int64_t m) {
  return m * (2 * lmax + 1 - m) / 2 + l;
}

}  // namespace ga

// ---- the plan (device tables for one (nside, lmax)) -----------------------
// Ring pair rp = 0 .. 2 nside - 1 pairs north ring rp (0-based, north to south)
// with south ring nrings - 1 - rp; rp = 2 nside - 1 is the equator (no partner).
struct ga_plan {
  int32_t nside, lmax, nrings, npairs;
  int64_t npix, nalm;
  int device;
  ga::ga_tuning tune;  // this plan's kernel choices
  // per ring pair (device): cos theta and sin theta of the north ring
  double *d_z = nullptr, *d_sth = nullptr;
  // per ring (device): first pixel, pixel count, phi0 = 2 pi pnum / pden
  int64_t *d_ring_start = nullptr;
  int32_t *d_ring_n = nullptr, *d_ring_pnum = nullptr, *d_ring_pden = nullptr;
  // Legendre: recurrence coefficients (a, b) per (l, m) in alm order (+4 pad),
  // lambda_l = a (x lambda_{l-1} - b lambda_{l-2}); seed constants c_m with
  // lambda_mm = c_m sin^m theta
  double2 *d_ab = nullptr;
  double2 *d_aab = nullptr;  // (a, a b) per (l, m): the same recurrence, one FMA deep
  // The split-FP16 synthesis runs the recurrence on y_l = lambda_l / c_l with
  // c_m = c_{m+1} = 1, c_l = a_l b_l c_{l-2}, which makes the lambda_{l-2}
  // coefficient 1: y_l = (gam_l x) y_{l-1} - y_{l-2}, gam_l = a_l c_{l-1} / c_l
  // (gam at l = m is 0: one step from (-seed, 0) yields y_m = seed); its B
  // operand carries c_l a_lm (d_csc).  lam_exp16: the power of two the
  // recurrence values are carried with, the largest <= 8 that keeps
  // max_l sqrt((2l+1)/4pi) / c_l times it below 2^15.5 (FP16 range).
  double *d_gam = nullptr;
  double *d_csc = nullptr;
  int lam_exp16 = 8;
  double *d_cm = nullptr;
  // FP16 MFMA stage: l-slabs of kSlab rows per m; slab_off[m] = index of the
  // first slab of m (lmax + 2 entries, slab_off[lmax + 1] = total)
  int64_t *d_slab_off = nullptr;
  int64_t nslab = 0;
  // Where the split-FP16 synthesis starts each wave's slab loop.  A wave
  // runs 64 consecutive ring pairs (group g = rp / 64) for one m; the leading
  // slabs in which every lane stays below the FP16 range contribute nothing.
  // Per (m, g): the first slab that is not dead for the whole wave
  // (d_f16_s0[m * f16_ngroups + g]); per (m, ring pair, padded to 64 per
  // group): the recurrence state (y_{l-2}, y_{l-1}) and scale count there
  // (d_f16_st / d_f16_k[m * 64 f16_ngroups + rp]).  Built on the device at
  // plan creation by the same code the contraction ran before its loop.
  int32_t f16_ngroups = 0;
  int32_t *d_f16_s0 = nullptr;
  double2 *d_f16_st = nullptr;
  int32_t *d_f16_k = nullptr;
  // spin recurrence coefficients (A, B, C) per (l, m) for s = 1, 2 (lazily built)
  double *d_spin[3] = {nullptr, nullptr, nullptr};
  // spin seed constants per m: [+s: (signed mantissa, exponent)] then [-s: ...]
  double2 *d_spin_seed[3] = {nullptr, nullptr, nullptr};
  // ring FFT: sub-length r per ring (n = 4 r), Bluestein tables
  // per distinct non-power-of-two r: chirp exp(+i pi k^2 / r) (r entries) and
  // the bit-reversed transform of its conjugate, scaled by 1 / L (L entries).
  int32_t max_r = 0, max_L = 0;
  int64_t *d_blu_off = nullptr;  // per r (0..max_r): offset into chirp tables, -1 if pow2
  double2 *d_blu64 = nullptr;
  float2 *d_blu32 = nullptr;
  // rings by FFT class (power-of-two r / Bluestein), with the class maxima
  int32_t *d_rings_pow2 = nullptr, *d_rings_blu = nullptr;
  int32_t n_pow2 = 0, n_blu = 0, max_r_pow2 = 0, max_r_blu = 0, max_L_blu = 0;
  // the power-of-two rings split once more: "wide" rings (r = r_wide, the
  // largest power-of-two r, with 4 r > lmax, so every bin has at most one
  // direct and one conjugate term; 256 <= r_wide <= 4096) and the rest
  int32_t *d_rings_wide = nullptr, *d_rings_narrow = nullptr;
  int32_t n_wide = 0, n_narrow = 0, r_wide = 0, max_r_narrow = 0;
  // the Bluestein rings split the same way: wide (4 r > lmax, r <= 2048,
  // L <= 4096) and the rest, with the class maxima
  int32_t *d_rings_bwide = nullptr, *d_rings_bnarrow = nullptr;
  int32_t n_bwide = 0, n_bnarrow = 0, max_r_bwide = 0, max_L_bwide = 0;
  int32_t max_r_bnarrow = 0, max_L_bnarrow = 0;
  // twiddles exp(2 pi i j / tw_len), j < tw_len; tw_len = max(max_L, 4 max_r_pow2)
  // (a power of two), so the radix-4 split of a power-of-two ring reads its
  // twiddles from the table too
  int32_t tw_len = 0;
  double2 *d_tw64 = nullptr;
  float2 *d_tw32 = nullptr;
  size_t table_bytes = 0;
};
