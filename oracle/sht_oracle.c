/*
 * FP64 CPU oracle for the HEALPix spherical-harmonic transforms on the GLASS
 * field-generation path.  TEST INFRASTRUCTURE ONLY: this file is compiled into
 * oracle/_build/libsht_oracle.so and is loaded solely by tests/, by the smoke
 * check and by the CPU-baseline leg of bench.py.  The product (glass_amd/)
 * never links or calls it.
 *
 * What it restates.  GLASS delegates the transforms to healpy
 * (glass/healpix.py:71 alm2map, :107 alm2map_spin, :270 map2alm), whose
 * compiled backend is not part of the reference tree.  This file restates the
 * published algorithm of those calls in plain FP64 C:
 *   - HEALPix RING geometry (Gorski et al. 2005, ApJ 622, 759, section 4);
 *   - fully normalised associated Legendre functions lambda_lm(theta) with the
 *     Condon-Shortley phase, by the standard three-term recurrence in l, with
 *     explicit binary-exponent tracking so that sin^m(theta) seeds never
 *     underflow;
 *   - spin-weighted lambda_lm for s = 1, 2 (Goldberg et al. 1967 seeds, Wigner-d
 *     three-term recurrence), combined in the HEALPix E/B convention
 *     Q + iU = -sum (E + iB) sY_lm;
 *   - per-ring Fourier sums by an exact complex DFT (radix-2 FFT, Bluestein for
 *     other lengths), with the ring phase phi0 reduced exactly in integers;
 *   - analysis as the adjoint with uniform pixel weight 4 pi / npix plus
 *     healpy's Jacobi refinement (alm += A(map - S(alm)), `iter` times).
 *
 * alm layout is healpy's: index(l, m) = m (2 lmax + 1 - m) / 2 + l, with
 * mmax = lmax, complex128 stored as interleaved (re, im) doubles.
 * Maps are RING ordered float64 of length 12 nside^2.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
  int64_t start; /* first pixel of the ring */
  int64_t n;     /* pixels in the ring */
  double z, sth; /* cos(theta), sin(theta) */
  int64_t pden;  /* phi0 = 2 pi * pnum / pden */
  int64_t pnum;
} ring_t;

/* Ring r is 1-based, 1 .. 4 nside - 1, north to south. */
static void ring_info(int64_t nside, int64_t r, ring_t *o) {
  const int64_t npix = 12 * nside * nside;
  const int64_t ncap = 2 * nside * (nside - 1);
  if (r < nside) {
    double t = (double)(r * r) / (3.0 * (double)(nside * nside));
    o->n = 4 * r;
    o->start = 2 * r * (r - 1);
    o->z = 1.0 - t;
    o->sth = sqrt(t * (2.0 - t));
    o->pden = 8 * r;
    o->pnum = 1;
  } else if (r <= 3 * nside) {
    o->n = 4 * nside;
    o->start = ncap + (r - nside) * 4 * nside;
    o->z = (double)(2 * nside - r) * 2.0 / (3.0 * (double)nside);
    o->sth = sqrt((1.0 - o->z) * (1.0 + o->z));
    o->pden = 8 * nside;
    o->pnum = ((r + nside) & 1) ? 0 : 1;
  } else {
    int64_t i = 4 * nside - r;
    double t = (double)(i * i) / (3.0 * (double)(nside * nside));
    o->n = 4 * i;
    o->start = npix - 2 * i * (i + 1);
    o->z = -(1.0 - t);
    o->sth = sqrt(t * (2.0 - t));
    o->pden = 8 * i;
    o->pnum = 1;
  }
}

void oracle_ring_info(int64_t nside, int64_t r, int64_t *start, int64_t *n,
                      double *z, double *sth, double *phi0) {
  ring_t o;
  ring_info(nside, r, &o);
  *start = o.start;
  *n = o.n;
  *z = o.z;
  *sth = o.sth;
  *phi0 = 2.0 * M_PI * (double)o.pnum / (double)o.pden;
}

static inline int64_t alm_index(int64_t lmax, int64_t l, int64_t m) {
  return m * (2 * lmax + 1 - m) / 2 + l;
}

/* ------------------------------------------------------------------------ */
/* complex DFT: x_j = sum_k a_k exp(sign * 2 pi i j k / n)                   */

static void fft_pow2(double *re, double *im, int64_t n, int sign) {
  /* iterative radix-2, bit-reversed input permutation */
  for (int64_t i = 1, j = 0; i < n; ++i) {
    int64_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double t = re[i]; re[i] = re[j]; re[j] = t;
      t = im[i]; im[i] = im[j]; im[j] = t;
    }
  }
  for (int64_t len = 2; len <= n; len <<= 1) {
    int64_t h = len >> 1;
    for (int64_t k = 0; k < h; ++k) {
      double ang = sign * 2.0 * M_PI * (double)k / (double)len;
      double wr = cos(ang), wi = sin(ang);
      for (int64_t s = 0; s < n; s += len) {
        int64_t a = s + k, b = a + h;
        double tr = re[b] * wr - im[b] * wi;
        double ti = re[b] * wi + im[b] * wr;
        re[b] = re[a] - tr; im[b] = im[a] - ti;
        re[a] += tr; im[a] += ti;
      }
    }
  }
}

/* any n >= 1: radix-2 when possible, Bluestein chirp-z otherwise */
static void dft_any(double *re, double *im, int64_t n, int sign) {
  if (n <= 1) return;
  if ((n & (n - 1)) == 0) {
    fft_pow2(re, im, n, sign);
    return;
  }
  int64_t L = 1;
  while (L < 2 * n - 1) L <<= 1;
  double *ar = calloc((size_t)L, sizeof(double)), *ai = calloc((size_t)L, sizeof(double));
  double *br = calloc((size_t)L, sizeof(double)), *bi = calloc((size_t)L, sizeof(double));
  double *cr = calloc((size_t)n, sizeof(double)), *ci = calloc((size_t)n, sizeof(double));
  /* chirp c_k = exp(sign * i pi k^2 / n), k^2 reduced mod 2n exactly */
  for (int64_t k = 0; k < n; ++k) {
    int64_t q = (int64_t)(((__int128)k * k) % (2 * n));
    double ang = sign * M_PI * (double)q / (double)n;
    cr[k] = cos(ang);
    ci[k] = sin(ang);
  }
  /* jk = (j^2 + k^2 - (j-k)^2)/2:  x_j = c_j sum_k (a_k c_k) conj(c_{j-k}) */
  for (int64_t k = 0; k < n; ++k) {
    ar[k] = re[k] * cr[k] - im[k] * ci[k];
    ai[k] = re[k] * ci[k] + im[k] * cr[k];
  }
  br[0] = cr[0]; bi[0] = -ci[0];
  for (int64_t k = 1; k < n; ++k) {
    br[k] = br[L - k] = cr[k];
    bi[k] = bi[L - k] = -ci[k];
  }
  fft_pow2(ar, ai, L, -1);
  fft_pow2(br, bi, L, -1);
  for (int64_t k = 0; k < L; ++k) {
    double tr = ar[k] * br[k] - ai[k] * bi[k];
    double ti = ar[k] * bi[k] + ai[k] * br[k];
    ar[k] = tr; ai[k] = ti;
  }
  fft_pow2(ar, ai, L, +1);
  for (int64_t j = 0; j < n; ++j) {
    double tr = ar[j] / (double)L, ti = ai[j] / (double)L;
    re[j] = tr * cr[j] - ti * ci[j];
    im[j] = tr * ci[j] + ti * cr[j];
  }
  free(ar); free(ai); free(br); free(bi); free(cr); free(ci);
}

void oracle_dft(double *re, double *im, int64_t n, int sign) { dft_any(re, im, n, sign); }

/* ring phase exp(i m phi0) with m phi0 reduced exactly */
static inline void ring_phase(const ring_t *g, int64_t m, double *c, double *s) {
  int64_t q = (m * g->pnum) % g->pden;
  double ang = 2.0 * M_PI * (double)q / (double)g->pden;
  *c = cos(ang);
  *s = sin(ang);
}

/* f_j = sum_{m=-mmax}^{mmax} F_m exp(i m phi_j), F_{-m} = conj(F_m), real out */
static void ring_synth(const ring_t *g, int64_t mmax, const double *Fr,
                       const double *Fi, double *out, double *wr, double *wi) {
  int64_t n = g->n;
  memset(wr, 0, (size_t)n * sizeof(double));
  memset(wi, 0, (size_t)n * sizeof(double));
  for (int64_t m = 0; m <= mmax; ++m) {
    double c, s;
    ring_phase(g, m, &c, &s);
    double ar = Fr[m] * c - Fi[m] * s, ai = Fr[m] * s + Fi[m] * c;
    int64_t k = m % n;
    wr[k] += ar; wi[k] += ai;
    if (m > 0) {
      int64_t k2 = (n - k) % n;
      wr[k2] += ar; wi[k2] -= ai;
    }
  }
  dft_any(wr, wi, n, +1);
  for (int64_t j = 0; j < n; ++j) out[j] = wr[j];
}

/* F_m = w exp(-i m phi0) sum_j f_j exp(-2 pi i j m / n), m = 0..mmax */
static void ring_anal(const ring_t *g, int64_t mmax, const double *f, double w,
                      double *Fr, double *Fi, double *wr, double *wi) {
  int64_t n = g->n;
  for (int64_t j = 0; j < n; ++j) { wr[j] = f[j]; wi[j] = 0.0; }
  dft_any(wr, wi, n, -1);
  for (int64_t m = 0; m <= mmax; ++m) {
    double c, s;
    ring_phase(g, m, &c, &s);
    int64_t k = m % n;
    /* (wr + i wi)(c - i s) */
    Fr[m] = w * (wr[k] * c + wi[k] * s);
    Fi[m] = w * (wi[k] * c - wr[k] * s);
  }
}

/* ------------------------------------------------------------------------ */
/* scalar Legendre: value = v * 2^e, rescaled so v never overflows          */

typedef struct { double v; int64_t e; } scaled_t;

static inline void sc_norm(scaled_t *a) {
  int ex;
  if (a->v == 0.0) { a->e = 0; return; }
  a->v = frexp(a->v, &ex);
  a->e += ex;
}

/* recurrence coefficients for one m: lambda_l = a_l (x lambda_{l-1} - b_l lambda_{l-2}) */
static void legendre_coeffs(int64_t m, int64_t lmax, double *ab) {
  for (int64_t l = m + 1; l <= lmax; ++l) {
    double ll = (double)l, mm = (double)m;
    ab[2 * (l - m)] = sqrt((4.0 * ll * ll - 1.0) / (ll * ll - mm * mm));
    ab[2 * (l - m) + 1] = sqrt(((ll - 1.0) * (ll - 1.0) - mm * mm) /
                               (4.0 * (ll - 1.0) * (ll - 1.0) - 1.0));
  }
}

/* coefficient table for all m, laid out like the alm (2 doubles per (l,m)) */
static double *legendre_table(int64_t lmax) {
  const int64_t nalm = (lmax + 1) * (lmax + 2) / 2;
  double *t = calloc((size_t)(2 * nalm), sizeof(double));
  for (int64_t m = 0; m <= lmax; ++m)
    legendre_coeffs(m, lmax, t + 2 * (m * (2 * lmax + 1 - m) / 2 + m));
  return t;
}

/* lambda_lm(x) for l = m..lmax into out[l-m]; seed (v,e) = lambda_mm */
static void legendre_column(int64_t m, int64_t lmax, double x, scaled_t seed,
                            const double *ab, double *out) {
  double p0 = 0.0, p1 = seed.v; /* lambda_{l-1}, lambda_l, shared exponent */
  int64_t e = seed.e;
  const double big = 0x1p300, rescale = 0x1p-600;
  for (int64_t l = m; l <= lmax; ++l) {
    if (l > m) {
      double pn = ab[2 * (l - m)] * (x * p1 - ab[2 * (l - m) + 1] * p0);
      p0 = p1;
      p1 = pn;
      if (fabs(p1) > big) { p1 *= rescale; p0 *= rescale; e += 600; }
    }
    out[l - m] = (e == 0) ? p1 : ((e < -1100) ? 0.0 : ldexp(p1, (int)e));
  }
}

/* lambda_mm seeds for m = 0..mmax at sin(theta) = sth, as scaled values */
static void legendre_seeds(int64_t mmax, double sth, scaled_t *seeds) {
  scaled_t s = {1.0 / sqrt(4.0 * M_PI), 0};
  sc_norm(&s);
  seeds[0] = s;
  for (int64_t m = 1; m <= mmax; ++m) {
    s.v *= -sqrt((2.0 * m + 1.0) / (2.0 * m)) * sth;
    sc_norm(&s);
    seeds[m] = s;
  }
}

/* exposed for tests: lambda_lm(cos theta) for one (m, theta), l = m..lmax */
void oracle_legendre(int64_t m, int64_t lmax, double z, double sth, double *out) {
  scaled_t *seeds = malloc((size_t)(m + 1) * sizeof(scaled_t));
  double *ab = calloc((size_t)(2 * (lmax - m + 1)), sizeof(double));
  legendre_seeds(m, sth, seeds);
  legendre_coeffs(m, lmax, ab);
  legendre_column(m, lmax, z, seeds[m], ab, out);
  free(seeds);
  free(ab);
}

/* ------------------------------------------------------------------------ */
/* spin-weighted lambda: out_p = sY, out_m = -sY, l = l0..lmax, l0=max(m,s)  */

static double lfact(int64_t n) { return lgamma((double)n + 1.0); }

static void spin_column(int64_t s, int64_t m, int64_t lmax, double z, double sth,
                        double *outp, double *outm) {
  int64_t l0 = m > s ? m : s;
  /* half-angle functions, accurate at both poles */
  double sh = sqrt(0.5 * (1.0 - z)), ch = sqrt(0.5 * (1.0 + z));
  if (z > 0.5) sh = sth / (2.0 * ch);
  else if (z < -0.5) ch = sth / (2.0 * sh);
  int64_t dif = m > s ? m - s : s - m;
  /* Goldberg seeds:  +s: sin^{m+s}(t/2) cos^{|m-s|}(t/2) (-1)^m
                      -s: sin^{|m-s|}(t/2) cos^{m+s}(t/2) (-1)^m (m>=s), (-1)^s (m<s) */
  double lnorm = 0.5 * (log((2.0 * l0 + 1.0) / (4.0 * M_PI)) + lfact(2 * l0) -
                        lfact(m + s) - lfact(dif));
  double lsh = log(sh), lch = log(ch);
  double lp = lnorm + (double)(m + s) * lsh + (double)dif * lch;
  double lm = lnorm + (double)dif * lsh + (double)(m + s) * lch;
  double sgp = (m & 1) ? -1.0 : 1.0;
  double sgm = (m >= s) ? sgp : ((s & 1) ? -1.0 : 1.0);
  /* carry log-magnitude as an exponent of 2 */
  const double LN2 = 0.69314718055994530942;
  int64_t ep = (int64_t)floor(lp / LN2), em = (int64_t)floor(lm / LN2);
  double vp = sgp * exp(lp - (double)ep * LN2), vm = sgm * exp(lm - (double)em * LN2);
  double pp0 = 0.0, pp1 = vp, pm0 = 0.0, pm1 = vm;
  const double big = 0x1p300, rescale = 0x1p-600;
  for (int64_t l = l0; l <= lmax; ++l) {
    if (l > l0) {
      double L = (double)(l - 1); /* recurrence from (L-1, L) to L+1 = l */
      double mm = (double)m, ss = (double)s;
      double den = sqrt(((L + 1) * (L + 1) - mm * mm) * ((L + 1) * (L + 1) - ss * ss));
      double A = sqrt((2 * L + 3) * (2 * L + 1)) * (L + 1) / den;
      double B = mm * ss / (L * (L + 1));
      double C = sqrt(2 * L + 3) * (L + 1) * sqrt((L * L - mm * mm) * (L * L - ss * ss)) /
                 (L * sqrt(2 * L - 1) * den);
      double np = A * (z + B) * pp1 - C * pp0;
      double nm = A * (z - B) * pm1 - C * pm0;
      pp0 = pp1; pp1 = np; pm0 = pm1; pm1 = nm;
      if (fabs(pp1) > big) { pp1 *= rescale; pp0 *= rescale; ep += 600; }
      if (fabs(pm1) > big) { pm1 *= rescale; pm0 *= rescale; em += 600; }
    }
    outp[l - l0] = (ep < -1100) ? 0.0 : ldexp(pp1, (int)ep);
    outm[l - l0] = (em < -1100) ? 0.0 : ldexp(pm1, (int)em);
  }
}

void oracle_spin_legendre(int64_t s, int64_t m, int64_t lmax, double z, double sth,
                          double *outp, double *outm) {
  spin_column(s, m, lmax, z, sth, outp, outm);
}

/* ------------------------------------------------------------------------ */

static int64_t nrings_of(int64_t nside) { return 4 * nside - 1; }

/* F_m of OR_BLK rings at once: independent recurrence chains in lockstep
 * (first with per-ring rescaling while any ring is below range, then plain) */
#define OR_BLK 8
static void synth_block(int64_t nside, int64_t m, int64_t lmax, const int64_t *rings, int nb,
                        const scaled_t *seeds, int64_t seed_stride, const double *ab,
                        const double *a, double *sr, double *si) {
  double x[OR_BLK], p0[OR_BLK], p1[OR_BLK], ar[OR_BLK], ai[OR_BLK];
  int64_t e[OR_BLK];
  for (int j = 0; j < OR_BLK; ++j) {
    ring_t g;
    ring_info(nside, rings[j < nb ? j : 0], &g);
    scaled_t sd = seeds[(size_t)(j < nb ? j : 0) * (size_t)seed_stride + (size_t)m];
    x[j] = g.z; p0[j] = 0.0; p1[j] = sd.v; e[j] = sd.e;
    ar[j] = 0.0; ai[j] = 0.0;
  }
  const double big = 0x1p300, rescale = 0x1p-600;
  int64_t l = m;
  int any_low = 1;
  for (; l <= lmax && any_low; ++l) {
    const double c0 = ab[2 * (l - m)], c1 = ab[2 * (l - m) + 1];
    any_low = 0;
    for (int j = 0; j < OR_BLK; ++j) {
      if (l > m) {
        double pn = c0 * (x[j] * p1[j] - c1 * p0[j]);
        p0[j] = p1[j]; p1[j] = pn;
      }
      if (e[j] < 0 && fabs(p1[j]) > big) { p1[j] *= rescale; p0[j] *= rescale; e[j] += 600; }
      double v = (e[j] == 0) ? p1[j] : ((e[j] < -1100) ? 0.0 : ldexp(p1[j], (int)e[j]));
      if (e[j] != 0) any_low = 1;
      ar[j] += v * a[2 * (l - m)];
      ai[j] += v * a[2 * (l - m) + 1];
    }
  }
  for (; l <= lmax; ++l) {
    const double c0 = ab[2 * (l - m)], c1 = ab[2 * (l - m) + 1];
    const double br = a[2 * (l - m)], bi = a[2 * (l - m) + 1];
    for (int j = 0; j < OR_BLK; ++j) {
      double pn = c0 * (x[j] * p1[j] - c1 * p0[j]);
      p0[j] = p1[j]; p1[j] = pn;
      ar[j] += pn * br;
      ai[j] += pn * bi;
    }
  }
  for (int j = 0; j < nb; ++j) { sr[j] = ar[j]; si[j] = ai[j]; }
}

/* lambda_mm seeds of every (ring, m) for the rings in the list */
static scaled_t *seed_table(int64_t nside, int64_t lmax, const int64_t *rings, int64_t nsel) {
  scaled_t *t = malloc((size_t)nsel * (size_t)(lmax + 1) * sizeof(scaled_t));
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int64_t i = 0; i < nsel; ++i) {
    ring_t g;
    ring_info(nside, rings[i], &g);
    legendre_seeds(lmax, g.sth, t + (size_t)i * (size_t)(lmax + 1));
  }
  return t;
}

/* alm (complex interleaved, healpy order, mmax = lmax) -> map (RING), rings
 * r_first, r_first + r_step, ... only (1-based; r_first = 1, r_step = 1 is the
 * full map).  Loop order m-outer so that one m's coefficients and alm stay in
 * cache while every ring is visited: F[ring][m] first, then the ring sums. */
void oracle_alm2map_rings(int64_t nside, int64_t lmax, const double *alm, double *map,
                          int64_t r_first, int64_t r_step, int nthreads) {
  const int64_t nr = nrings_of(nside);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  int64_t nsel = 0;
  for (int64_t r = r_first; r <= nr; r += r_step) ++nsel;
  int64_t *rings = malloc((size_t)(nsel > 0 ? nsel : 1) * sizeof(int64_t));
  nsel = 0;
  for (int64_t r = r_first; r <= nr; r += r_step) rings[nsel++] = r;
  double *tab = legendre_table(lmax);
  scaled_t *seeds = seed_table(nside, lmax, rings, nsel);
  double *F = malloc((size_t)nsel * (size_t)(lmax + 1) * 2 * sizeof(double));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int64_t m = 0; m <= lmax; ++m) {
    const double *ab = tab + 2 * alm_index(lmax, m, m);
    const double *a = alm + 2 * alm_index(lmax, m, m);
    for (int64_t i0 = 0; i0 < nsel; i0 += OR_BLK) {
      const int nb = (int)((nsel - i0) < OR_BLK ? (nsel - i0) : OR_BLK);
      double sr[OR_BLK], si[OR_BLK];
      synth_block(nside, m, lmax, rings + i0, nb,
                  seeds + (size_t)i0 * (size_t)(lmax + 1), lmax + 1, ab, a, sr, si);
      for (int j = 0; j < nb; ++j) {
        double *f = F + 2 * ((size_t)(i0 + j) * (size_t)(lmax + 1) + (size_t)m);
        f[0] = sr[j];
        f[1] = (m == 0) ? 0.0 : si[j]; /* healpy ignores Im a_l0 */
      }
    }
  }
#ifdef _OPENMP
#pragma omp parallel
#endif
  {
    double *Fr = malloc((size_t)(lmax + 1) * sizeof(double));
    double *Fi = malloc((size_t)(lmax + 1) * sizeof(double));
    double *wr = malloc((size_t)(16 * nside) * sizeof(double));
    double *wi = malloc((size_t)(16 * nside) * sizeof(double));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
    for (int64_t i = 0; i < nsel; ++i) {
      ring_t g;
      ring_info(nside, rings[i], &g);
      for (int64_t m = 0; m <= lmax; ++m) {
        Fr[m] = F[2 * ((size_t)i * (size_t)(lmax + 1) + (size_t)m)];
        Fi[m] = F[2 * ((size_t)i * (size_t)(lmax + 1) + (size_t)m) + 1];
      }
      ring_synth(&g, lmax, Fr, Fi, map + g.start, wr, wi);
    }
    free(Fr); free(Fi); free(wr); free(wi);
  }
  free(F); free(seeds); free(tab); free(rings);
}

/* The same synthesis with the equatorial symmetry a tuned CPU SHT uses (the
 * CPU-baseline form): one recurrence per (m, north ring) serves the ring and
 * its south mirror, lambda_lm(-x) = (-1)^(l+m) lambda_lm(x), so the l-sums are
 * split by the parity of l - m into E and O and F_north = E + O,
 * F_south = E - O.  A ring's value is carried as p 2^e while it is below
 * 2^-900 (where it cannot change an FP64 sum of order one and is left out);
 * the scale is folded in as soon as it is above, and from the l where every
 * ring of the block is folded the recurrence runs unscaled, two l at a time. */
#define SYM_BLK 16
static void synth_block_sym(int64_t m, int64_t lmax, const double *x, const scaled_t *sd,
                            int nb, const double *ab, const double *a, double *fn, double *fs) {
  double p0[SYM_BLK], p1[SYM_BLK], er[SYM_BLK], ei[SYM_BLK], orr[SYM_BLK], oi[SYM_BLK];
  int64_t e[SYM_BLK];
  for (int j = 0; j < SYM_BLK; ++j) {
    const int jj = j < nb ? j : 0;
    p0[j] = 0.0; p1[j] = sd[jj].v; e[j] = sd[jj].e;
    er[j] = ei[j] = orr[j] = oi[j] = 0.0;
  }
  const double big = 0x1p300, rescale = 0x1p-600;
  int64_t l = m;
  int any_low = 1;
  for (; l <= lmax && any_low; ++l) {
    const double c0 = ab[2 * (l - m)], c1 = ab[2 * (l - m) + 1];
    const int odd = (int)((l - m) & 1);
    any_low = 0;
    for (int j = 0; j < SYM_BLK; ++j) {
      if (l > m) {
        double pn = c0 * (x[j] * p1[j] - c1 * p0[j]);
        p0[j] = p1[j]; p1[j] = pn;
      }
      if (e[j] != 0) {
        if (fabs(p1[j]) > big) { p1[j] *= rescale; p0[j] *= rescale; e[j] += 600; }
        /* fold the scale in once the value is well inside the normal range */
        if (p1[j] != 0.0 && ilogb(p1[j]) + e[j] > -900) {
          p1[j] = ldexp(p1[j], (int)e[j]);
          p0[j] = ldexp(p0[j], (int)e[j]);
          e[j] = 0;
        } else {
          any_low = 1; /* below 2^-900: contributes nothing at FP64 */
          continue;
        }
      }
      const double v = p1[j];
      if (odd) { orr[j] += v * a[2 * (l - m)]; oi[j] += v * a[2 * (l - m) + 1]; }
      else { er[j] += v * a[2 * (l - m)]; ei[j] += v * a[2 * (l - m) + 1]; }
    }
  }
  /* plain phase, two steps (one even, one odd l - m) at a time */
  if (l <= lmax && ((l - m) & 1)) {
    const double c0 = ab[2 * (l - m)], c1 = ab[2 * (l - m) + 1];
    for (int j = 0; j < SYM_BLK; ++j) {
      double pn = c0 * (x[j] * p1[j] - c1 * p0[j]);
      p0[j] = p1[j]; p1[j] = pn;
      orr[j] += pn * a[2 * (l - m)]; oi[j] += pn * a[2 * (l - m) + 1];
    }
    ++l;
  }
  for (; l + 1 <= lmax; l += 2) {
    const double c0 = ab[2 * (l - m)], c1 = ab[2 * (l - m) + 1];
    const double d0 = ab[2 * (l + 1 - m)], d1 = ab[2 * (l + 1 - m) + 1];
    const double br = a[2 * (l - m)], bi = a[2 * (l - m) + 1];
    const double cr = a[2 * (l + 1 - m)], ci = a[2 * (l + 1 - m) + 1];
    for (int j = 0; j < SYM_BLK; ++j) {
      const double pe = c0 * (x[j] * p1[j] - c1 * p0[j]);
      const double po = d0 * (x[j] * pe - d1 * p1[j]);
      p0[j] = pe; p1[j] = po;
      er[j] += pe * br; ei[j] += pe * bi;
      orr[j] += po * cr; oi[j] += po * ci;
    }
  }
  if (l <= lmax) {
    const double c0 = ab[2 * (l - m)], c1 = ab[2 * (l - m) + 1];
    for (int j = 0; j < SYM_BLK; ++j) {
      const double pe = c0 * (x[j] * p1[j] - c1 * p0[j]);
      er[j] += pe * a[2 * (l - m)]; ei[j] += pe * a[2 * (l - m) + 1];
    }
  }
  for (int j = 0; j < nb; ++j) {
    fn[2 * j] = er[j] + orr[j]; fn[2 * j + 1] = ei[j] + oi[j];
    fs[2 * j] = er[j] - orr[j]; fs[2 * j + 1] = ei[j] - oi[j];
  }
}

void oracle_alm2map_sym(int64_t nside, int64_t lmax, const double *alm, double *map,
                        int nthreads) {
  const int64_t nr = nrings_of(nside), nnorth = 2 * nside; /* rings 1 .. 2 nside */
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  int64_t *rings = malloc((size_t)nnorth * sizeof(int64_t));
  double *x = malloc((size_t)nnorth * sizeof(double));
  for (int64_t i = 0; i < nnorth; ++i) {
    ring_t g;
    rings[i] = i + 1;
    ring_info(nside, i + 1, &g);
    x[i] = g.z;
  }
  double *tab = legendre_table(lmax);
  scaled_t *seeds = seed_table(nside, lmax, rings, nnorth);
  /* F[ring - 1][m], all 4 nside - 1 rings */
  double *F = malloc((size_t)nr * (size_t)(lmax + 1) * 2 * sizeof(double));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int64_t m = 0; m <= lmax; ++m) {
    const double *ab = tab + 2 * alm_index(lmax, m, m);
    const double *a = alm + 2 * alm_index(lmax, m, m);
    for (int64_t i0 = 0; i0 < nnorth; i0 += SYM_BLK) {
      const int nb = (int)((nnorth - i0) < SYM_BLK ? (nnorth - i0) : SYM_BLK);
      double xb[SYM_BLK], fn[2 * SYM_BLK], fs[2 * SYM_BLK];
      scaled_t sd[SYM_BLK];
      for (int j = 0; j < SYM_BLK; ++j) {
        const int64_t i = i0 + (j < nb ? j : 0);
        xb[j] = x[i];
        sd[j] = seeds[(size_t)i * (size_t)(lmax + 1) + (size_t)m];
      }
      synth_block_sym(m, lmax, xb, sd, nb, ab, a, fn, fs);
      for (int j = 0; j < nb; ++j) {
        const int64_t rn = i0 + j + 1, rs = 4 * nside - rn;
        double *f = F + 2 * ((size_t)(rn - 1) * (size_t)(lmax + 1) + (size_t)m);
        f[0] = fn[2 * j];
        f[1] = (m == 0) ? 0.0 : fn[2 * j + 1];
        if (rs != rn) {
          f = F + 2 * ((size_t)(rs - 1) * (size_t)(lmax + 1) + (size_t)m);
          f[0] = fs[2 * j];
          f[1] = (m == 0) ? 0.0 : fs[2 * j + 1];
        }
      }
    }
  }
#ifdef _OPENMP
#pragma omp parallel
#endif
  {
    double *Fr = malloc((size_t)(lmax + 1) * sizeof(double));
    double *Fi = malloc((size_t)(lmax + 1) * sizeof(double));
    double *wr = malloc((size_t)(16 * nside) * sizeof(double));
    double *wi = malloc((size_t)(16 * nside) * sizeof(double));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
    for (int64_t r = 1; r <= nr; ++r) {
      ring_t g;
      ring_info(nside, r, &g);
      for (int64_t m = 0; m <= lmax; ++m) {
        Fr[m] = F[2 * ((size_t)(r - 1) * (size_t)(lmax + 1) + (size_t)m)];
        Fi[m] = F[2 * ((size_t)(r - 1) * (size_t)(lmax + 1) + (size_t)m) + 1];
      }
      ring_synth(&g, lmax, Fr, Fi, map + g.start, wr, wi);
    }
    free(Fr); free(Fi); free(wr); free(wi);
  }
  free(F); free(seeds); free(tab); free(x); free(rings);
}

void oracle_alm2map(int64_t nside, int64_t lmax, const double *alm, double *map,
                    int nthreads) {
  oracle_alm2map_rings(nside, lmax, alm, map, 1, 1, nthreads);
}

/* adjoint of synthesis with weight 4 pi / npix per pixel (one pass):
 * ring transforms first, then per m (each m owns its alm column) */
static void alm_from_map_once(int64_t nside, int64_t lmax, const double *map,
                              double *alm, int nthreads) {
  const int64_t nr = nrings_of(nside);
  const double w = 4.0 * M_PI / (double)(12 * nside * nside);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  int64_t *rings = malloc((size_t)nr * sizeof(int64_t));
  for (int64_t i = 0; i < nr; ++i) rings[i] = i + 1;
  double *tab = legendre_table(lmax);
  scaled_t *seeds = seed_table(nside, lmax, rings, nr);
  double *F = malloc((size_t)nr * (size_t)(lmax + 1) * 2 * sizeof(double));
#ifdef _OPENMP
#pragma omp parallel
#endif
  {
    double *Fr = malloc((size_t)(lmax + 1) * sizeof(double));
    double *Fi = malloc((size_t)(lmax + 1) * sizeof(double));
    double *wr = malloc((size_t)(16 * nside) * sizeof(double));
    double *wi = malloc((size_t)(16 * nside) * sizeof(double));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
    for (int64_t i = 0; i < nr; ++i) {
      ring_t g;
      ring_info(nside, rings[i], &g);
      ring_anal(&g, lmax, map + g.start, w, Fr, Fi, wr, wi);
      for (int64_t m = 0; m <= lmax; ++m) {
        F[2 * ((size_t)i * (size_t)(lmax + 1) + (size_t)m)] = Fr[m];
        F[2 * ((size_t)i * (size_t)(lmax + 1) + (size_t)m) + 1] = Fi[m];
      }
    }
    free(Fr); free(Fi); free(wr); free(wi);
  }
#ifdef _OPENMP
#pragma omp parallel
#endif
  {
    double *lam = malloc((size_t)(lmax + 1) * sizeof(double));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int64_t m = 0; m <= lmax; ++m) {
      const double *ab = tab + 2 * alm_index(lmax, m, m);
      double *a = alm + 2 * alm_index(lmax, m, m);
      for (int64_t l = m; l <= lmax; ++l) { a[2 * (l - m)] = 0.0; a[2 * (l - m) + 1] = 0.0; }
      for (int64_t i = 0; i < nr; ++i) {
        ring_t g;
        ring_info(nside, rings[i], &g);
        legendre_column(m, lmax, g.z, seeds[(size_t)i * (size_t)(lmax + 1) + (size_t)m], ab, lam);
        const double fr = F[2 * ((size_t)i * (size_t)(lmax + 1) + (size_t)m)];
        const double fi = F[2 * ((size_t)i * (size_t)(lmax + 1) + (size_t)m) + 1];
        for (int64_t l = m; l <= lmax; ++l) {
          a[2 * (l - m)] += lam[l - m] * fr;
          a[2 * (l - m) + 1] += lam[l - m] * fi;
        }
      }
      if (m == 0)
        for (int64_t l = 0; l <= lmax; ++l) a[2 * l + 1] = 0.0; /* m = 0 real */
    }
    free(lam);
  }
  free(F); free(seeds); free(tab); free(rings);
}

/* healpy map2alm(iter=niter): alm = A f; repeat alm += A(f - S alm) */
void oracle_map2alm(int64_t nside, int64_t lmax, const double *map, double *alm,
                    int niter, int nthreads) {
  const int64_t npix = 12 * nside * nside;
  const int64_t nalm = (lmax + 1) * (lmax + 2) / 2;
  alm_from_map_once(nside, lmax, map, alm, nthreads);
  if (niter <= 0) return;
  double *res = malloc((size_t)npix * sizeof(double));
  double *da = malloc((size_t)(2 * nalm) * sizeof(double));
  for (int it = 0; it < niter; ++it) {
    oracle_alm2map(nside, lmax, alm, res, nthreads);
    for (int64_t p = 0; p < npix; ++p) res[p] = map[p] - res[p];
    alm_from_map_once(nside, lmax, res, da, nthreads);
    for (int64_t i = 0; i < 2 * nalm; ++i) alm[i] += da[i];
  }
  free(res);
  free(da);
}

/* spin-s synthesis: [E, B] alm -> [Q, U] maps, Q + iU = -sum (E + iB) sY_lm */
void oracle_alm2map_spin(int64_t nside, int64_t lmax, int64_t spin, const double *almE,
                         const double *almB, double *mapQ, double *mapU, int nthreads) {
  const int64_t nr = nrings_of(nside);
  const double sgn = (spin & 1) ? -1.0 : 1.0; /* (-1)^s */
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    double *Qr = malloc((size_t)(lmax + 1) * sizeof(double));
    double *Qi = malloc((size_t)(lmax + 1) * sizeof(double));
    double *Ur = malloc((size_t)(lmax + 1) * sizeof(double));
    double *Ui = malloc((size_t)(lmax + 1) * sizeof(double));
    double *lp = malloc((size_t)(lmax + 1) * sizeof(double));
    double *lm = malloc((size_t)(lmax + 1) * sizeof(double));
    double *wr = malloc((size_t)(16 * nside) * sizeof(double));
    double *wi = malloc((size_t)(16 * nside) * sizeof(double));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
    for (int64_t r = 1; r <= nr; ++r) {
      ring_t g;
      ring_info(nside, r, &g);
      for (int64_t m = 0; m <= lmax; ++m) {
        double qr = 0, qi = 0, ur = 0, ui = 0;
        int64_t l0 = m > spin ? m : spin;
        if (l0 <= lmax) {
          spin_column(spin, m, lmax, g.z, g.sth, lp, lm);
          const double *e = almE + 2 * alm_index(lmax, m, m);
          const double *b = almB + 2 * alm_index(lmax, m, m);
          for (int64_t l = l0; l <= lmax; ++l) {
            double fp = -0.5 * (lp[l - l0] + sgn * lm[l - l0]);
            double fm = -0.5 * (lp[l - l0] - sgn * lm[l - l0]);
            double er = e[2 * (l - m)], ei = (m == 0) ? 0.0 : e[2 * (l - m) + 1];
            double br = b[2 * (l - m)], bi = (m == 0) ? 0.0 : b[2 * (l - m) + 1];
            /* Q_m = E F+ + i B F-,  U_m = B F+ - i E F- */
            qr += er * fp - bi * fm;
            qi += ei * fp + br * fm;
            ur += br * fp + ei * fm;
            ui += bi * fp - er * fm;
          }
        }
        Qr[m] = qr; Qi[m] = qi; Ur[m] = ur; Ui[m] = ui;
      }
      ring_synth(&g, lmax, Qr, Qi, mapQ + g.start, wr, wi);
      ring_synth(&g, lmax, Ur, Ui, mapU + g.start, wr, wi);
    }
    free(Qr); free(Qi); free(Ur); free(Ui); free(lp); free(lm); free(wr); free(wi);
  }
}
