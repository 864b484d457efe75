// Host-side launchers for the HIP kernels (internal to the library).
#pragma once

#include "common.h"

namespace ga {

// Legendre stage, FP64 plain: alm (healpy order complex128) -> F [ring][m] complex128
int launch_legendre_synth_f64(const ga_plan *p, const double *alm, double *F, hipStream_t s);

// Legendre stage, FP32 MFMA batched: alm [idx][S] complex64 -> F (grouped ring buffer)
int launch_legendre_synth_mfma(const ga_plan *p, const ga_tuning &tu, int S, const float *alm,
                               float *F, hipStream_t s);
// Legendre stage, split-FP16 MFMA batched: same in/out, in two launches:
// operand preparation (alm -> split, swizzled B operands and column
// exponents: the bws / colexp parts of the FP32 workspace, f32_workspace) and
// the contraction proper
// (rows l < lmin of the B operands are zero)
// (colpart: nparts rows of per-block column maxima standing in for the alm
// rows in the column-maximum pass; csc: per-(l, m) factors applied to the
// rows, the c_l of the scalar synthesis's normalised recurrence, NULL for none)
int launch_legendre_prepare_f16(const ga_plan *p, int S, const float *alm, void *bws,
                                int32_t *colexp, hipStream_t s, int lmin = 0,
                                const float *colpart = nullptr, int64_t nparts = 0,
                                const double *csc = nullptr);
// nt: column tiles per wave (0 automatic, 1, 2, 4 = wave pairs); flush: the
// accumulation form (f16_accumulation: -1 one chain, n > 0 register master
// sums every n slabs, -2 two l-chunks with the first parked in F and the O
// half of its equator row at park ((lmax + 1) 2 S floats), -3 two l-chunks
// with the first parked at park, one 16 nt KiB slot per wave of the grid)
// mgroup: consecutive m per XCD turn (1 or 4); chunk_min: least l-range in
// slabs that the two-chunk form splits (0: kTwoChunkMinSlabs)
int launch_legendre_contract_f16(const ga_plan *p, int S, float *F, const void *bws,
                                 const int32_t *colexp, hipStream_t s, int nt, int flush,
                                 void *park, int mgroup, int chunk_min);
// fills the plan's start table of the split-FP16 synthesis (d_f16_s0,
// d_f16_st, d_f16_k, allocated by the caller) from its recurrence tables
int launch_legendre_f16_start(const ga_plan *p, hipStream_t s);
// the batched Legendre stage on the path chosen by the tuning set tu (a
// snapshot of the plan's, the same for the preparation and the contraction
// of one synthesis); ws is the FP32 workspace with F at its start
// (prepare + contract = synth, which snapshots p->tune once)
int launch_legendre_prepare_batch(const ga_plan *p, const ga_tuning &tu, int S, const float *alm,
                                  void *ws, hipStream_t s, const float *colpart = nullptr,
                                  int64_t nparts = 0);
int launch_legendre_contract_batch(const ga_plan *p, const ga_tuning &tu, int S,
                                   const float *alm, void *ws, hipStream_t s);
int launch_legendre_synth_batch(const ga_plan *p, int S, const float *alm, void *ws,
                                hipStream_t s);
int launch_legendre_synth_batch(const ga_plan *p, const ga_tuning tu, int S, const float *alm,
                                void *ws, hipStream_t s);
// the split-FP16 contraction of up to kGroupedMax geometries (plans[i], S[i]
// shells, prepared FP32 workspace ws[i]) in one grid, each on its plan's
// tiles and accumulation form
constexpr int kGroupedMax = 8;
int launch_legendre_contract_grouped(int ngroup, const ga_plan *const *plans, const int32_t *S,
                                     void *const *ws, hipStream_t s);

// Ring stage of synthesis: F -> maps [S][npix] with fused transform.  F holds
// the shells in groups of G: offset ((ring ngrp + s/G)(lmax+1) + m) G + s%G
// (G = S is the plain [ring][m][S] layout; the MFMA stages write
// G = ring_group(S)).
int launch_ring_synth_f64(const ga_plan *p, int S, int G, const double *F, double *maps,
                          const ga_epilogue *epi, hipStream_t s);
int launch_ring_synth_f32(const ga_plan *p, int S, int G, const float *F, float *maps,
                          const ga_epilogue *epi, hipStream_t s);

// Ring stage of analysis: maps [S][npix] -> F [ring][m][S] (weighted, phase removed)
int launch_ring_anal_f64(const ga_plan *p, int S, const double *maps, double *F, hipStream_t s);
int launch_ring_anal_f32(const ga_plan *p, int S, const float *maps, float *F, hipStream_t s);
// out = ref - FP32 ring synthesis of F (fused-pass forms only; GA_ERR_UNSUPPORTED
// with nothing launched otherwise)
int launch_ring_resid_f32(const ga_plan *p, int S, int G, const float *F, float *out,
                          const float *ref, hipStream_t s);

// Legendre stage of analysis: F [ring][m][S] -> alm [idx][S] (+= when accumulate)
// partial: anal_partial_slots(p) * nalm * S complex scratch
int anal_partial_slots(const ga_plan *p);
int launch_legendre_anal_f64(const ga_plan *p, int S, const double *F, double *alm,
                             double *partial, bool accumulate, hipStream_t s);
int launch_legendre_anal_f32(const ga_plan *p, int S, const float *F, float *alm, float *partial,
                             bool accumulate, hipStream_t s);
// the split-FP16 MFMA form of the FP32 analysis contraction: F -> partial
// sums [slot][idx][S] complex, one slot per 512 ring pairs (the caller adds them)
int launch_legendre_anal_f16(const ga_plan *p, int S, const float *F, float *partial,
                             hipStream_t s);

// Spin synthesis Legendre stage: (E, B) [idx][S] -> ring buffer [ring][m][2S],
// column 2s = Q_s, 2s + 1 = U_s
int launch_spin_synth_f64(const ga_plan *p, int spin, int S, const double *E, const double *B,
                          double *Fo, hipStream_t s);
int launch_spin_synth_f32(const ga_plan *p, int spin, int S, const float *E, const float *B,
                          float *Fo, hipStream_t s);
// Spin synthesis Legendre stage on split-FP16 MFMA: E (and B, may be NULL for
// B = 0) [idx][S] complex64 -> the grouped ring buffer of 2S transforms (Q_s =
// transform 2s, U_s = 2s + 1; groups of ring_group(2S)) at the start of the FP32
// workspace ws of 2S transforms (f32_workspace(p, 2S))
int launch_spin_synth_f16(const ga_plan *p, int spin, int S, const float *E, const float *B,
                          void *ws, hipStream_t s);

int plan_spin_tables(ga_plan *p, int s);

}  // namespace ga
