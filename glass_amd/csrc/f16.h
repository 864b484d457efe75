// Split-FP16 MFMA pieces shared by the scalar and spin Legendre stages (gfx950).
//
// An FP32 operand x is carried as x = x_hi + x_lo, two FP16 numbers (22
// significant bits), and a product as three v_mfma_f32_32x32x16_f16:
//     lambda a ~ lambda_hi a_hi + lambda_hi a_lo + lambda_lo a_hi.
// Operand layouts of v_mfma_f32_32x32x16_f16 (wave64): lane L holds
// A[L % 32][8 (L / 32) + j] and B[8 (L / 32) + j][L % 32], j = 0..7; the
// accumulator element e of lane L is D[8 (e / 4) + 4 (L / 32) + e % 4][L % 32].
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ga {
namespace f16 {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

constexpr int kLamExp = 8;  // recurrence values are carried times 2^8

__device__ __forceinline__ uint32_t bits_h2(f16x2 h) { return __builtin_bit_cast(uint32_t, h); }

// a - (the low / high FP16 half of h, widened): one v_fma_mix_f32 (-h * 1 + a)
// with the half selected by op_sel; exact, since a - h is representable
__device__ __forceinline__ float sub_half_lo(float a, uint32_t h) {
  float r;
  asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h), "v"(a));
  return r;
}
__device__ __forceinline__ float sub_half_hi(float a, uint32_t h) {
  float r;
  asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h),
      "v"(a));
  return r;
}

// x = hi + lo for two values at once: hi rounded toward zero (one packed
// conversion; its error is captured exactly by the residual), lo = the FP32
// residual rounded to nearest, formed and rounded to FP16 by v_fma_mixlo_f16 /
// v_fma_mixhi_f16 (a * 1 - h in FP32, written to one half of lo): three
// instructions per pair.
__device__ __forceinline__ void split_pair(float a, float b, uint32_t &hi, uint32_t &lo) {
  const auto h = __builtin_amdgcn_cvt_pkrtz(a, b);
  hi = __builtin_bit_cast(uint32_t, h);
  uint32_t l;  // the high half written by mixlo is garbage until mixhi overwrites it
  asm("v_fma_mixlo_f16 %0, %1, 1.0, -%2 op_sel_hi:[0,0,1]" : "=v"(l) : "v"(a), "v"(hi));
  asm("v_fma_mixhi_f16 %0, %1, 1.0, -%2 op_sel:[0,0,1] op_sel_hi:[0,0,1]" : "+v"(l)
      : "v"(b), "v"(hi));
  lo = l;
}

// A operands of one 32-row slab from a lane's packed values.  ph / pl
// [parity][group] hold the lane's 32 values of the slab as packed FP16 pairs
// (group g: k = 2g, 2g + 1 of the 16 rows of that parity);
// v_permlane32_swap(X, Y) = ([X_lo | Y_lo], [X_hi | Y_hi]) turns them into
// the A operands of two row tiles: tile 0 = the rows computed by lanes 0..31,
// tile 1 = those of lanes 32..63.  ah / al [parity][row tile].
__device__ __forceinline__ void slab_a(const uint32_t (&ph)[2][8], const uint32_t (&pl)[2][8],
                                       uint4 (&ah)[2][2], uint4 (&al)[2][2]) {
  uint32_t h[2][2][4], l[2][2][4];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const auto sh = __builtin_amdgcn_permlane32_swap(ph[p][w], ph[p][w + 4], false, false);
      const auto sl = __builtin_amdgcn_permlane32_swap(pl[p][w], pl[p][w + 4], false, false);
      h[p][0][w] = sh[0];
      h[p][1][w] = sh[1];
      l[p][0][w] = sl[0];
      l[p][1][w] = sl[1];
    }
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      ah[p][t] = uint4{h[p][t][0], h[p][t][1], h[p][t][2], h[p][t][3]};
      al[p][t] = uint4{l[p][t][0], l[p][t][1], l[p][t][2], l[p][t][3]};
    }
}

// The MFMAs of one slab: acc [parity][row tile][column tile] += A B with the
// three split products; bh / bl [parity][column tile] are the pre-swizzled B
// operands.  Accumulators are interleaved so that dependent MFMAs are
// 4 NT - 1 instructions apart.
template <int NT>
__device__ __forceinline__ void slab_mfma_a(const uint4 (&ah)[2][2], const uint4 (&al)[2][2],
                                            const uint4 (&bh)[2][NT], const uint4 (&bl)[2][NT],
                                            f32x16 (&acc)[2][2][NT]) {
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) {
          const f16x8 a = __builtin_bit_cast(f16x8, (q == 2) ? al[p][t] : ah[p][t]);
          const f16x8 bb = __builtin_bit_cast(f16x8, (q == 1) ? bl[p][ct] : bh[p][ct]);
          acc[p][t][ct] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, bb, acc[p][t][ct], 0, 0, 0);
        }
}

// both steps
template <int NT>
__device__ __forceinline__ void slab_mfma(const uint32_t (&ph)[2][8], const uint32_t (&pl)[2][8],
                                          const uint4 (&bh)[2][NT], const uint4 (&bl)[2][NT],
                                          f32x16 (&acc)[2][2][NT]) {
  uint4 ah[2][2], al[2][2];
  slab_a(ph, pl, ah, al);
  slab_mfma_a<NT>(ah, al, bh, bl, acc);
}

}  // namespace f16
}  // namespace ga
