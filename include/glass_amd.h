/*
 * glass_amd: MI355X-native field-generation hot path for GLASS.
 *
 * Plain C ABI.  Every array argument is a raw device pointer owned by the
 * caller; sizes are element counts; the stream is a hipStream_t passed as
 * void*.  Every entry point returns a status code (GA_OK == 0) and leaves a
 * human-readable message for ga_last_error() on failure.  Work is
 * stream-ordered: a call returns after enqueueing (except where noted).
 *
 * Layouts
 *   alm, single map : complex, healpy order idx(l,m) = m(2 lmax + 1 - m)/2 + l,
 *                     mmax = lmax, interleaved (re, im).
 *   alm, batch      : complex64, [idx(l,m)][S]  (shell index innermost).
 *   maps            : RING ordered, 12 nside^2 pixels; batches are [S][npix].
 *   ring workspace  : complex [nrings][lmax+1][S], nrings = 4 nside - 1.
 *
 * Each entry point names the reference interface it stands in for.
 */
#ifndef GLASS_AMD_H
#define GLASS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum ga_status {
  GA_OK = 0,
  GA_ERR_INVALID = 1,     /* bad argument (maps to ValueError) */
  GA_ERR_HIP = 2,         /* HIP runtime error */
  GA_ERR_NOMEM = 3,       /* device allocation failed (plan tables) */
  GA_ERR_UNSUPPORTED = 4, /* size outside what the kernels handle */
  GA_ERR_NOT_POSDEF = 5,  /* "covariance matrix is not positive definite" */
  GA_ERR_WORKSPACE = 6    /* workspace smaller than ga_*_workspace_size */
};

/* pointwise transformation fused into the synthesis epilogue;
 * glass/grf/_transformations.py:26-28 (Normal), :83-89 (Lognormal),
 * :170-175 (SquaredNormal) */
enum ga_transform { GA_T_NORMAL = 0, GA_T_LOGNORMAL = 1, GA_T_SQUARED_NORMAL = 2 };

typedef struct ga_epilogue {
  int32_t kind;   /* enum ga_transform */
  int32_t pad;
  double lamda;   /* lambda (shift / scale) */
  double param;   /* Lognormal: variance of the Gaussian field; SquaredNormal: a */
} ga_epilogue;

typedef struct ga_plan ga_plan;

/* Tuning knobs: which kernel form each batched FP32 stage runs (chosen by
 * the autotuner; no counterpart in the reference, which has no device code).
 * Each plan carries its own set, copied at ga_plan_create from the process
 * default; ga_plan_set_tuning changes one plan's set (a set-up step: not
 * concurrently with launches on that plan), and every entry point reads its
 * plan's set once per call.  ga_set_tuning / ga_get_tuning read and write the
 * process default, which only affects plans created afterwards. */
enum ga_tune_key {
  GA_TUNE_MFMA_TILES = 1, /* 32-column tiles per wave: 1, 2, 0 = automatic, or 4 =
                             producer / consumer wave pairs (split FP16 only) */
  GA_TUNE_MFMA_CHUNK = 2, /* FP32 path: quads per accumulation chunk, 0 = none */
  GA_TUNE_MFMA_PATH = 3,  /* 0 = automatic (split FP16), 1 = FP32 MFMA, 2 = split FP16 */
  GA_TUNE_RING_PATH = 4,  /* FP32 ring stage: 0 = automatic (default: synthesis
                             by the register-fold / register-split kernels, analysis
                             by the register form), 1 = radix-4 LDS stages (both),
                             2 = synthesis by four-array fused passes on every ring */
  GA_TUNE_SPIN_PATH = 5,  /* FP32 spin synthesis: 0 = automatic (split FP16 MFMA),
                             1 = VALU, 2 = split FP16 MFMA */
  GA_TUNE_ANAL_PATH = 6,  /* FP32 analysis Legendre stage: 0 = automatic (split FP16
                             MFMA), 1 = VALU, 2 = split FP16 MFMA */
  GA_TUNE_F16_FLUSH = 7   /* split-FP16 synthesis accumulation (shorter FP32 rounding
                             chains): n > 0 = every n slabs of 32 l rows the MFMA
                             accumulators are added into FP32 master sums in
                             registers (one column tile per wave); -3 / -2 = the
                             l-range cut in two chunks, the first parked and added
                             back at the end: -3 in a region of the workspace as
                             large as the ring buffer (coalesced, fastest), -2 in the
                             ring buffer itself (the workspace grows by only
                             (lmax + 1) 2 nshell floats; needs 64 ring rows of the
                             batch within 4 GiB); size the workspace under the same
                             setting; -1 = one chain; 0 = automatic (8 when
                             lmax > 4096, else -3) */,
  GA_TUNE_M_GROUP = 8     /* split-FP16 synthesis block order: the m values of an XCD
                             come in runs of this many consecutive m (1 or 4; 0 =
                             automatic = 4), so that with 4 the four m of a 128-byte
                             line of the ring buffer are written by one XCD's L2
                             (bit-identical results; 37.4 -> 33.9 ms at nside 2048,
                             64 shells with the ring buffer in groups of 4 shells;
                             no measurable difference with groups of 8) */,
  GA_TUNE_CHUNK_MIN = 9   /* two-chunk accumulation (GA_TUNE_F16_FLUSH -2 / -3): the
                             least l-range, in slabs of 32 rows, that is split (1 ..
                             4096; 0 = automatic = 64: shorter chains add little to
                             the rounding error) */
};
int ga_set_tuning(int32_t key, int32_t value);
int ga_get_tuning(int32_t key, int32_t *value);

/* Last error message of the calling host thread ("" if none). */
const char *ga_last_error(void);

/* Library version string. */
const char *ga_version(void);

/* HEALPix geometry and Legendre/FFT tables for one (nside, lmax) on the current
 * device.  Allocates the plan's own read-only tables; no allocation happens in
 * any other entry point.  One of the tables (where each wave of the split-FP16
 * synthesis starts its slab loop, and its recurrence state there) is computed
 * by a kernel on the null stream, which this call waits for: create plans
 * outside stream capture.  Replaces the per-call set-up inside healpy.alm2map /
 * healpy.map2alm (glass/healpix.py:71, :270). */
int ga_plan_create(int32_t nside, int32_t lmax, ga_plan **out);
int ga_plan_destroy(ga_plan *plan);
/* One plan's tuning knobs (see enum ga_tune_key). */
int ga_plan_set_tuning(ga_plan *plan, int32_t key, int32_t value);
int ga_plan_get_tuning(const ga_plan *plan, int32_t key, int32_t *value);
/* Bytes of device memory held by the plan's tables. */
int ga_plan_table_bytes(const ga_plan *plan, size_t *bytes);

/* Workspace bytes for nbatch maps at the given element width (4 = fp32,
 * 8 = fp64).  Synthesis (scalar): the ring buffer [nrings][lmax+1][nbatch]
 * complex.  Spin synthesis needs the size for 2 nbatch.  Analysis: use
 * ga_map2alm_workspace_size. */
int ga_workspace_size(const ga_plan *plan, int32_t nbatch, int32_t real_bytes,
                      size_t *bytes);
int ga_map2alm_workspace_size(const ga_plan *plan, int32_t nbatch, int32_t real_bytes,
                              size_t *bytes);

/* One-time build of the spin-s (s = 1, 2) recurrence tables of a plan
 * (allocates); required before the spin entry points. */
int ga_plan_prepare_spin(ga_plan *plan, int32_t spin);

/* FP64 scalar synthesis of one map, plain (no MFMA, no batching).
 * glass/healpix.py:38-78 alm2map(alms, nside, pol=False), as called at
 * glass/fields.py:429 and glass/lensing.py:326. */
int ga_alm2map_f64(const ga_plan *plan, const double *alm, double *map,
                   void *workspace, size_t ws_bytes, void *stream);

/* FP32 batched scalar synthesis of S maps: Legendre contraction on FP32 MFMA
 * with the recurrence in FP64, per-ring FFT, fused pointwise transform
 * (epi: S device structs, or NULL for identity).  Batched form of the
 * per-shell alm2map + t(x, var) of glass/fields.py:429, :894. */
int ga_alm2map_batch_f32(const ga_plan *plan, int32_t nshell, const float *alm,
                         float *maps, const ga_epilogue *epi, void *workspace,
                         size_t ws_bytes, void *stream);

/* ga_alm2map_batch_f32 without a workspace argument and without any device
 * allocation: the workspace is a fixed scratch array in the library's device
 * code (ga_noalloc_scratch_bytes: 128 MiB per device, held from the library's
 * load until the process ends), for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, hosts that forbid hipMalloc).
 * It runs the two-chunk accumulation in its compact form (the first chunk
 * parked in the ring buffer, GA_TUNE_F16_FLUSH = -2) where the plan would
 * use the form with a region of its own (-3): the maps are bit-identical.
 * Sizes whose workspace under that form exceeds the scratch are rejected
 * with GA_ERR_UNSUPPORTED before anything is launched;
 * the current device must be the plan's.  Every call on one GPU shares that
 * GPU's scratch, which keeps the last call's contents: calls that overlap on
 * one GPU (different streams, or host threads) race and must be serialised by
 * the caller.  No counterpart in the reference. */
int ga_noalloc_scratch_bytes(size_t *bytes);
int ga_alm2map_batch_f32_noalloc(const ga_plan *plan, int32_t nshell, const float *alm,
                                 float *maps, const ga_epilogue *epi, void *stream);

/* The two stages of ga_alm2map_batch_f32, callable separately (for timing
 * and for callers that keep the ring buffer): the MFMA Legendre stage
 * alm [idx][S] -> ring buffer F, which occupies the start of the workspace
 * (ga_workspace_size(plan, S, 4) bytes; shells in groups of G = 8 when
 * S >= 8, else G = 4, complex element (ring, m, s) at
 * ((ring ceil(S/G) + s/G)(lmax+1) + m) G + s%G), and
 * the ring-FFT stage F (= the workspace pointer) -> maps [S][npix] with the
 * fused transform. */
int ga_legendre_synth_batch_f32(const ga_plan *plan, int32_t nshell, const float *alm,
                                void *workspace, size_t ws_bytes, void *stream);
/* ga_legendre_synth_batch_f32 in its two launches: operand preparation (on
 * the split-FP16 path: per-column scaling, FP16 splitting and swizzling of the
 * alm into the workspace; nothing on the FP32 path) and the contraction into
 * F.  The contraction reads what the preparation of the same nshell and alm
 * left in the workspace, under the same tuning. */
int ga_legendre_prepare_batch_f32(const ga_plan *plan, int32_t nshell, const float *alm,
                                  void *workspace, size_t ws_bytes, void *stream);
int ga_legendre_contract_batch_f32(const ga_plan *plan, int32_t nshell, const float *alm,
                                   void *workspace, size_t ws_bytes, void *stream);
/* ga_legendre_prepare_batch_f32 with the per-column maxima taken from nparts
 * rows of per-block maxima (colpart, [nparts][2 nshell] floats, as written by
 * ga_fields_alm_colmax for the same alm) instead of from the nalm alm rows;
 * the workspace it leaves is identical.  No counterpart in the reference:
 * it saves the batched pipeline one full read of the alm. */
int ga_legendre_prepare_colmax_batch_f32(const ga_plan *plan, int32_t nshell, const float *alm,
                                         const float *colpart, int64_t nparts, void *workspace,
                                         size_t ws_bytes, void *stream);
int ga_ring_synth_batch_f32(const ga_plan *plan, int32_t nshell, const float *F, float *maps,
                            const ga_epilogue *epi, void *stream);

/* The split-FP16 contraction of an uneven batch in ONE launch: ngroup
 * (1..8) geometries, each a plan, its shell count nshell[i] and its FP32
 * workspace (ws_bytes[i] >= ga_workspace_size(plans[i], nshell[i], 4)) that
 * ga_legendre_prepare_batch_f32 has prepared for the same shells.  The
 * blocks of all geometries share one grid (list the largest first: its
 * blocks are dispatched first and the small geometries fill the device as
 * it drains); each geometry runs on its own plan's tiles and accumulation
 * form and writes its ring buffer F at the start of its workspace, exactly
 * as ga_legendre_contract_batch_f32 would.  Rejected (GA_ERR_UNSUPPORTED):
 * a plan tuned to the FP32 MFMA path or accumulating in register master
 * sums (lmax > 4096 by default).  The batched form of many alm2map calls of
 * different nside (glass/healpix.py:38-78 via glass/fields.py:429). */
int ga_legendre_contract_grouped_f32(int32_t ngroup, const ga_plan *const *plans,
                                     const int32_t *nshell, void *const *workspace,
                                     const size_t *ws_bytes, void *stream);

/* Graph capture for C hosts: every entry point of this library that takes a
 * workspace allocates nothing and sets no memory with memset, so a sequence
 * of calls on one stream can be recorded once and replayed with one host
 * call.  ga_graph_begin starts capturing `stream` (a created stream, not the
 * null stream; thread-local capture mode); ga_graph_end stops it and
 * instantiates the recorded work; ga_graph_launch replays it on a stream;
 * ga_graph_destroy frees it.  Plans must exist, and spin tables be prepared
 * (ga_plan_prepare_spin), before the capture begins; the buffers the
 * recorded calls name must outlive the graph.  If a call fails during the
 * capture, ga_graph_end still ends it and returns the error.  No counterpart
 * in the reference (which calls healpy once per map): the graph and
 * batch-schedule entry points of the batched path. */
typedef struct ga_graph ga_graph;
int ga_graph_begin(void *stream);
int ga_graph_end(void *stream, ga_graph **out);
int ga_graph_launch(const ga_graph *graph, void *stream);
int ga_graph_destroy(ga_graph *graph);

/* FP64 scalar analysis with healpy's Jacobi refinement (niter passes,
 * alm += A(map - S(alm))), uniform pixel weights 4 pi / npix.
 * glass/healpix.py:232-276 map2alm(pol=False), called at glass/lensing.py:306.
 * map_tmp: npix float64 scratch (used when niter > 0). */
int ga_map2alm_f64(const ga_plan *plan, const double *map, double *alm, int32_t niter,
                   double *map_tmp, void *workspace, size_t ws_bytes, void *stream);

/* FP64 scalar analysis with per-pixel quadrature weights (npix doubles that
 * sum to 4 pi, e.g. from ga_pixel_weights) in place of the uniform
 * 4 pi / npix, and healpy's Jacobi refinement with the same weights
 * (alm += A_w(map - S(alm)), niter times).  map_tmp (npix doubles) is
 * required.  glass/healpix.py:232-276 map2alm(use_pixel_weights=True), as
 * called at glass/lensing.py:306 and :408. */
int ga_map2alm_weighted_f64(const ga_plan *plan, const double *map, const double *weights,
                            double *alm, int32_t niter, double *map_tmp, void *workspace,
                            size_t ws_bytes, void *stream);

/* Pixel quadrature weights of the plan's nside, computed instead of read from
 * healpy's weight files (which map2alm(use_pixel_weights=True) of
 * glass/healpix.py:232-276 needs): the smallest correction of the uniform
 * weights 4 pi / npix for which sum_p w_p Y*_lm(p) = sqrt(4 pi) delta_l0
 * delta_m0 for every l <= the plan's lmax (which must be <= 2 nside), by
 * conjugate gradients on the plan's FP64 synthesis and analysis until the
 * residual falls below tol (relative) or maxiter iterations.  Writes npix
 * weights, the iteration count and the final relative residual (either may
 * be NULL).  Synchronises the stream (it reads inner products back each
 * iteration): a set-up call, not capturable. */
int ga_pixel_weights_workspace_size(const ga_plan *plan, size_t *bytes);
int ga_pixel_weights(const ga_plan *plan, double tol, int32_t maxiter, double *weights,
                     int32_t *iterations, double *residual, void *workspace, size_t ws_bytes,
                     void *stream);

/* ga_map2alm_batch_f32 with per-pixel quadrature weights (npix doubles, the
 * same for every map) in place of the uniform 4 pi / npix, also in the
 * Jacobi passes; map_tmp is required.  The batched form of
 * map2alm(use_pixel_weights=True) at glass/lensing.py:306 and :408. */
int ga_map2alm_weighted_batch_f32(const ga_plan *plan, int32_t nshell, const float *maps,
                                  const double *weights, float *alm, int32_t niter,
                                  float *map_tmp, void *workspace, size_t ws_bytes,
                                  void *stream);

/* FP32 batched analysis of S maps [S][npix] into [idx][S] alm (niter Jacobi
 * passes; the residual synthesis uses the FP32-MFMA path).
 * map_tmp: S npix float32 scratch. */
int ga_map2alm_batch_f32(const ga_plan *plan, int32_t nshell, const float *maps, float *alm,
                         int32_t niter, float *map_tmp, void *workspace, size_t ws_bytes,
                         void *stream);

/* Spin-s synthesis (s = 1, 2) of [E, B] into [Q, U], HEALPix convention
 * Q + iU = -sum (E + iB) sY_lm.  glass/healpix.py:81-108 alm2map_spin, called
 * at glass/lensing.py:343 (s = 1) and :366, :428 (s = 2).
 * f64: one map pair, maps = [Q | U] (2 npix).  f32: S pairs [S][2][npix];
 * almB may be NULL there (B = 0, the lensing case: one contraction instead of
 * two).  workspace: ga_workspace_size for 2 S maps. */
int ga_alm2map_spin_f64(const ga_plan *plan, int32_t spin, const double *almE,
                        const double *almB, double *maps, void *workspace, size_t ws_bytes,
                        void *stream);
int ga_alm2map_spin_batch_f32(const ga_plan *plan, int32_t spin, int32_t nshell,
                              const float *almE, const float *almB, float *maps, void *workspace,
                              size_t ws_bytes, void *stream);

/* alm[idx(l,m)][s] *= fl[l] (0 beyond nfl), in place or out of place, for
 * S interleaved maps; real_bytes 4 or 8.  glass/healpix.py:111-140 almxfl. */
int ga_almxfl(int32_t lmax, int32_t nshell, int32_t real_bytes, const void *alm,
              const double *fl, int32_t nfl, void *out, void *stream);

/* GLASS-order multalm: alm[l(l+1)/2 + m] * bl[l] (complex128 in/out, float64
 * bl).  glass/harmonics.py:17-47. */
int ga_multalm_f64(int32_t nl, const double *alm, const double *bl, double *out,
                   void *stream);

/* GLASS (l-major) -> healpy (m-major) alm order, complex128.
 * glass/fields.py:943-962. */
int ga_glass_to_healpix_alm_f64(int32_t nl, const double *alm_glass, double *alm_hp,
                                void *stream);

/* Batched conditional-Gaussian scaling vectors for nshell consecutive fields.
 * cov: [nshell][nl][k+1] float64 covariance rows as cls2cov yields them
 * (auto-spectrum first, then the correlations with the previous fields,
 * already halved).  state: [nl][kmax*kmax + 2 kmax + 1] float64 carried between
 * calls (ga_iternorm_state_doubles(kmax) per mode); it has size kprev on entry
 * (k <= kprev + 1) unless start = 1, which marks the first row of this call as
 * field 0 and initialises the state.  w: [nshell][nl][k+1] output scaling
 * vectors.  *flag (device int32, caller-zeroed) is set to 1 if any covariance
 * is not positive definite.  glass/fields.py:101-188 iternorm. */
int ga_iternorm_state_doubles(int32_t kmax, int32_t *n);
int ga_iternorm_batch(int32_t nl, int32_t kprev, int32_t k, int32_t kmax, int32_t nshell,
                      int32_t start, const double *cov, double *state, double *w,
                      int32_t *flag, void *stream);

/* Correlated alm for shells first_shell .. first_shell + nshell - 1:
 *   alm_t = m0fix(to_healpix(sum_j w_t[l][j + mis] z_{t - len + 1 + j}))
 * (glass/fields.py:401-425).  Normal vectors z are either
 *   - host-drawn (numpy PCG64 stream): zring = device ring buffer of nslot
 *     GLASS-order complex128 vectors, z_t in slot t % nslot (nslot >=
 *     nshell + k so that every shell of the call sees its window); or
 *   - counter-based Philox4x32-10 (seed, shell, index) when zring == NULL:
 *     key = seed, counter = (index, shell / 2), words 0-1 of the call for
 *     an even shell and 2-3 for an odd one, Box-Muller on each pair.
 * Output: batch layout [idx][ldo] complex64 (out_f32; shell s of the call in
 * column s, ldo >= nshell, 0 = nshell, so that several calls can fill the
 * column ranges of one shared batch) or, for nshell == 1, healpy-order
 * complex128 (out_f64); exactly one of the two is non-NULL. */
int ga_fields_alm(int32_t nl, int32_t k, int32_t first_shell, int32_t nshell,
                  const double *w, const double *zring, int32_t nslot, uint64_t seed,
                  float *out_f32, int32_t ldo, double *out_f64, void *stream);
/* ga_fields_alm with the device generator into out_f32 (k + 1 <= 8), which
 * also writes, for each block of 256 consecutive healpy indices, the maxima
 * of |re| and |im| per shell: colpart[b][2 s + 0 / 1] for b < *rows of
 * ga_fields_colmax_rows(nl).  No counterpart in the reference (input of
 * ga_legendre_prepare_colmax_batch_f32). */
int ga_fields_colmax_rows(int32_t nl, int64_t *rows);
int ga_fields_alm_colmax(int32_t nl, int32_t k, int32_t first_shell, int32_t nshell,
                         const double *w, uint64_t seed, float *out_f32, int32_t ldo,
                         float *colpart, void *stream);

/* HEALPix pixel window functions for l = 0 .. lmax, computed from their
 * definition (the harmonic power of each pixel's normalised top hat, averaged
 * over the pixels; spin s from the spin-s harmonics): pw0 the scalar window,
 * pws (spin 1 or 2, NULL for spin 0) the spin-s window, 0 for l < s.  float64
 * device outputs; nside <= 8192.  Stands in for healpy.pixwin, called at
 * glass/healpix.py:313-356 (pol = True: spin 2) and used by
 * glass/lensing.py:353-362, :414-419 (discretized = True). */
int ga_pixwin_workspace_size(int32_t nside, int32_t lmax, size_t *bytes);
int ga_pixwin(int32_t nside, int32_t lmax, int32_t spin, double *pw0, double *pws,
              void *workspace, size_t ws_bytes, void *stream);

/* Multi-plane convergence step on n-pixel maps (real_bytes 4 = float32,
 * 8 = float64): kappa3 = (1 - t) kappa1 + t kappa2 + f delta2, evaluated as
 * the reference's three in-place float64 array operations (same roundings;
 * float32 maps rounded back after each); delta2 NULL = 0 (before the first
 * mass plane); kappa3 may alias kappa1.  t and f come from the cosmology on
 * the host.  glass/lensing.py:511-586 MultiPlaneConvergence.add_plane. */
int ga_multi_plane_update(int64_t n, int32_t real_bytes, const void *kappa1, const void *kappa2,
                          const void *delta2, double t, double f, void *kappa3, void *stream);

/* Points from a density contrast (glass/points.py:443-540), one population
 * per call on npix-pixel RING maps.  Workspace: ga_points_workspace_size.
 * ga_points_counts: the expected count per pixel (glass/points.py:214-316)
 *   lam = max(0, (b(delta) [- mean] + 1) * scale [* vis]),
 * scale = ARCMIN2_SPHERE / npix * ngal, b(delta) by model 0 = delta (bias
 * None, or a bias model applied by the caller), 1 = linear (bias * delta,
 * glass/points.py:115-134), 2 = log-linear (expm1(bias * log1p(delta)),
 * :137-160); delta float32 or float64 (real_bytes), vis float64 or NULL; the
 * mean is subtracted when remove_monopole.  Writes lam (float64, or NULL)
 * and/or counts (int64, or NULL): Poisson counts from the counter-based
 * Philox4x32-10 generator keyed by (seed, population, pixel), with the total
 * count in *total (device int64) and the block offsets kept in the workspace
 * for ga_points_positions.
 * ga_points_scan: the same offsets and total for counts drawn elsewhere (the
 *   host generator of glass/points.py:319-348).
 * ga_points_positions: every point of the counts placed uniformly in area
 *   inside its pixel (lon in [0, 360), lat, degrees, float64, total points),
 *   uniforms keyed by (seed, population, point); the workspace must hold the
 *   offsets of the same counts.  Stands in for healpix.randang as called at
 *   glass/points.py:425-429 (glass/healpix.py:398-432). */
int ga_points_workspace_size(int64_t npix, size_t *bytes);
int ga_points_counts(int64_t npix, int32_t real_bytes, const void *delta, int32_t model,
                     double bias, int32_t remove_monopole, double scale, const double *vis,
                     uint64_t seed, int64_t population, double *lam, int64_t *counts,
                     int64_t *total, void *workspace, size_t ws_bytes, void *stream);
int ga_points_scan(int64_t npix, const int64_t *counts, int64_t *total, void *workspace,
                   size_t ws_bytes, void *stream);
int ga_points_positions(int32_t nside, const int64_t *counts, uint64_t seed, int64_t population,
                        double *lon, double *lat, void *workspace, size_t ws_bytes,
                        void *stream);

/* Angular power spectrum <-> correlation function on n Gauss-Legendre nodes
 * (x ascending in (-1, 1), weights w; float64 device arrays), exact for
 * l < n:  corr(x_k) = sum_l (2l+1)/(4 pi) cl_l P_l(x_k),
 *         cl_l = 2 pi sum_k w_k corr(x_k) P_l(x_k).
 * Stand in for transformcl.cltocorr / corrtocl as called by the Gaussian
 * spectra solver (glass/grf/_solver.py:96-140, glass/grf/_core.py:184). */
int ga_legendre_nodes(int32_t n, double *x, double *w, void *stream);
int ga_cl2corr(int32_t nl, const double *cl, int32_t n, const double *x, double *corr,
               void *stream);
int ga_corr2cl_workspace_size(int32_t n, int32_t nl, size_t *bytes);
int ga_corr2cl(int32_t n, const double *x, const double *w, const double *corr, int32_t nl,
               double *cl, void *workspace, size_t ws_bytes, void *stream);

/* Pointwise transforms on a float64 map (n pixels), out may equal x.
 * glass/grf/_transformations.py. */
int ga_transform_f64(int64_t n, const double *x, const ga_epilogue *epi_host, double *out,
                     void *stream);

#ifdef __cplusplus
}
#endif

#endif /* GLASS_AMD_H */
