// Plan creation: HEALPix RING geometry, Legendre recurrence tables and ring-FFT
// tables for one (nside, lmax), computed on the host in FP64 and uploaded once.
//
// Geometry follows Gorski et al. 2005 (ApJ 622, 759), section 4: north polar
// cap rings i < nside hold 4 i pixels at z = 1 - i^2 / (3 nside^2) with
// phi0 = pi / (4 i); equatorial rings nside <= i <= 3 nside hold 4 nside pixels
// at z = 4/3 - 2 i / (3 nside) with phi0 = pi / (4 nside) when i + nside is even
// and 0 otherwise; the south cap mirrors the north.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "common.h"
#include "kernels.h"

namespace ga {

static thread_local std::string g_err;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
}
void clear_error() { g_err.clear(); }

ga_tuning &tuning() {
  static ga_tuning t;
  return t;
}

namespace {

struct RingGeom {
  int64_t start;
  int32_t n, pnum, pden;
  double z, sth;
};

RingGeom ring_geom(int64_t nside, int64_t r /* 1-based */) {
  RingGeom g{};
  const int64_t npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
  if (r < nside) {
    double t = double(r * r) / (3.0 * double(nside * nside));
    g.n = int32_t(4 * r);
    g.start = 2 * r * (r - 1);
    g.z = 1.0 - t;
    g.sth = std::sqrt(t * (2.0 - t));
    g.pden = int32_t(8 * r);
    g.pnum = 1;
  } else if (r <= 3 * nside) {
    g.n = int32_t(4 * nside);
    g.start = ncap + (r - nside) * 4 * nside;
    g.z = double(2 * nside - r) * 2.0 / (3.0 * double(nside));
    g.sth = std::sqrt((1.0 - g.z) * (1.0 + g.z));
    g.pden = int32_t(8 * nside);
    g.pnum = ((r + nside) & 1) ? 0 : 1;
  } else {
    int64_t i = 4 * nside - r;
    double t = double(i * i) / (3.0 * double(nside * nside));
    g.n = int32_t(4 * i);
    g.start = npix - 2 * i * (i + 1);
    g.z = -(1.0 - t);
    g.sth = std::sqrt(t * (2.0 - t));
    g.pden = int32_t(8 * i);
    g.pnum = 1;
  }
  return g;
}

bool is_pow2(int64_t x) { return x > 0 && (x & (x - 1)) == 0; }

int64_t blu_len(int64_t r) {
  int64_t L = 1;
  while (L < 2 * r - 1) L <<= 1;
  return L;
}

unsigned bitrev(unsigned x, int bits) {
  unsigned y = 0;
  for (int i = 0; i < bits; ++i) { y = (y << 1) | (x & 1u); x >>= 1; }
  return y;
}

// in-place forward complex FFT (exp(-2 pi i jk / n)), n power of two
void host_fft(std::vector<std::complex<double>> &a) {
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    for (size_t k = 0; k < len / 2; ++k) {
      double ang = -2.0 * M_PI * double(k) / double(len);
      std::complex<double> w(std::cos(ang), std::sin(ang));
      for (size_t s = 0; s < n; s += len) {
        auto u = a[s + k], v = a[s + k + len / 2] * w;
        a[s + k] = u + v;
        a[s + k + len / 2] = u - v;
      }
    }
  }
}

template <typename T>
int alloc_table(T **dst, size_t bytes, size_t *acc) {
  hipError_t e = hipMalloc((void **)dst, bytes);
  if (e != hipSuccess) {
    *dst = nullptr;
    set_error("hipMalloc(%zu) for plan tables failed: %s", bytes, hipGetErrorString(e));
    return GA_ERR_NOMEM;
  }
  *acc += bytes;
  return GA_OK;
}

template <typename T>
int upload(T **dst, const void *src, size_t bytes, size_t *acc) {
  if (bytes == 0) return GA_OK;
  hipError_t e = hipMalloc((void **)dst, bytes);
  if (e != hipSuccess) {
    set_error("hipMalloc(%zu) for plan tables failed: %s", bytes, hipGetErrorString(e));
    return GA_ERR_NOMEM;
  }
  GA_CHECK_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
  *acc += bytes;
  return GA_OK;
}

}  // namespace

// Spin recurrence tables, built on first use: per (l, m) in alm order,
// (A, A B, C) with  sl_{l} = (A x + A B) sl_{l-1} - C sl_{l-2}  (+s) and
// (A x - A B) for -s, from the Wigner-d three-term recurrence.  The entries
// for m <= l <= l0 = max(m, s) are (0, 0, -1): such a step moves sl_{l-2}
// into place, so a recurrence that starts at l = m with the seed in the right
// slot (which depends on the parity of l0 - m) reaches sl_{l0} = seed.
int plan_spin_tables(ga_plan *p, int s) {
  if (p->d_spin[s]) return GA_OK;
  const int64_t lmax = p->lmax, nalm = p->nalm;
  std::vector<double> t(size_t(3 * (nalm + 4)), 0.0);
  std::vector<double2> seeds(size_t(2 * (lmax + 1)), double2{0.0, 0.0});
  for (int64_t m = 0; m <= lmax; ++m) {
    int64_t l0 = m > s ? m : s;
    // (0, 0, -1) at l = m .. l0: one step from (seed, 0) yields the seed itself
    for (int64_t l = m; l <= l0 && l <= lmax; ++l) {
      size_t i0 = size_t(3 * hp_index(lmax, l, m));
      t[i0] = 0.0; t[i0 + 1] = 0.0; t[i0 + 2] = -1.0;
    }
    // |seed constant| = sqrt((2 l0 + 1)/(4 pi) (2 l0)! / ((m + s)! |m - s|!)), as m 2^e
    {
      const int64_t dif = m > s ? m - s : s - m;
      long double l2 = 0.5L * (std::log2((2.0L * l0 + 1.0L) / (4.0L * (long double)M_PI)) +
                               (std::lgamma((long double)(2 * l0 + 1)) -
                                std::lgamma((long double)(m + s + 1)) -
                                std::lgamma((long double)(dif + 1))) /
                                   std::log(2.0L));
      long double e = std::floor(l2);
      double mant = double(std::exp2(l2 - e));
      double sp = (m & 1) ? -1.0 : 1.0;
      double sm = (m >= s) ? sp : ((s & 1) ? -1.0 : 1.0);
      seeds[size_t(m)] = double2{sp * mant, double(e)};
      seeds[size_t(lmax + 1 + m)] = double2{sm * mant, double(e)};
    }
    for (int64_t l = l0 + 1; l <= lmax; ++l) {
      double L = double(l - 1), mm = double(m), ss = double(s);
      double den = std::sqrt(((L + 1) * (L + 1) - mm * mm) * ((L + 1) * (L + 1) - ss * ss));
      double A = std::sqrt((2 * L + 3) * (2 * L + 1)) * (L + 1) / den;
      double B = mm * ss / (L * (L + 1));
      double C = std::sqrt(2 * L + 3) * (L + 1) * std::sqrt((L * L - mm * mm) * (L * L - ss * ss)) /
                 (L * std::sqrt(2 * L - 1) * den);
      size_t i = size_t(3 * hp_index(lmax, l, m));
      t[i] = A; t[i + 1] = A * B; t[i + 2] = C;
    }
  }
  int rc = upload(&p->d_spin[s], t.data(), t.size() * sizeof(double), &p->table_bytes);
  if (rc) return rc;
  return upload(&p->d_spin_seed[s], seeds.data(), seeds.size() * sizeof(double2), &p->table_bytes);
}

}  // namespace ga

using namespace ga;

extern "C" const char *ga_last_error(void) { return g_err.c_str(); }
extern "C" const char *ga_version(void) { return "glass_amd 0.1 (gfx950)"; }

namespace ga {
namespace {
// one knob of a tuning set, by key; value checks shared by the process
// default and the per-plan setters
int tuning_set(ga_tuning &t, int32_t key, int32_t value) {
  switch (key) {
    case GA_TUNE_MFMA_TILES:
      GA_REQUIRE(value == 0 || value == 1 || value == 2 || value == 4, GA_ERR_INVALID,
                 "mfma tiles must be 0, 1, 2 or 4");
      t.mfma_nt = value;
      return GA_OK;
    case GA_TUNE_MFMA_CHUNK:
      GA_REQUIRE(value >= 0 && value <= 4096, GA_ERR_INVALID, "mfma chunk out of range");
      t.mfma_chunk = value;
      return GA_OK;
    case GA_TUNE_MFMA_PATH:
      GA_REQUIRE(value >= 0 && value <= 2, GA_ERR_INVALID, "mfma path must be 0, 1 or 2");
      t.mfma_path = value;
      return GA_OK;
    case GA_TUNE_RING_PATH:
      GA_REQUIRE(value >= 0 && value <= 2, GA_ERR_INVALID, "ring path must be 0, 1 or 2");
      t.ring_path = value;
      return GA_OK;
    case GA_TUNE_SPIN_PATH:
      GA_REQUIRE(value >= 0 && value <= 2, GA_ERR_INVALID, "spin path must be 0, 1 or 2");
      t.spin_path = value;
      return GA_OK;
    case GA_TUNE_ANAL_PATH:
      GA_REQUIRE(value >= 0 && value <= 2, GA_ERR_INVALID, "analysis path must be 0, 1 or 2");
      t.anal_path = value;
      return GA_OK;
    case GA_TUNE_CHUNK_MIN:
      GA_REQUIRE(value >= 0 && value <= 4096, GA_ERR_INVALID, "chunk min out of range");
      t.chunk_min = value;
      return GA_OK;
    case GA_TUNE_M_GROUP:
      GA_REQUIRE(value == 0 || value == 1 || value == 4, GA_ERR_INVALID, "m group must be 0, 1 or 4");
      t.m_group = value;
      return GA_OK;
    case GA_TUNE_F16_FLUSH:
      GA_REQUIRE(value >= -3 && value <= 4096, GA_ERR_INVALID, "f16 flush out of range");
      t.f16_flush = value;
      return GA_OK;
    default:
      GA_REQUIRE(false, GA_ERR_INVALID, "unknown tuning key %d", key);
  }
}

int tuning_get(const ga_tuning &t, int32_t key, int32_t *value) {
  GA_REQUIRE(value, GA_ERR_INVALID, "NULL value");
  switch (key) {
    case GA_TUNE_MFMA_TILES: *value = t.mfma_nt; return GA_OK;
    case GA_TUNE_MFMA_CHUNK: *value = t.mfma_chunk; return GA_OK;
    case GA_TUNE_MFMA_PATH: *value = t.mfma_path; return GA_OK;
    case GA_TUNE_RING_PATH: *value = t.ring_path; return GA_OK;
    case GA_TUNE_SPIN_PATH: *value = t.spin_path; return GA_OK;
    case GA_TUNE_ANAL_PATH: *value = t.anal_path; return GA_OK;
    case GA_TUNE_F16_FLUSH: *value = t.f16_flush; return GA_OK;
    case GA_TUNE_M_GROUP: *value = t.m_group; return GA_OK;
    case GA_TUNE_CHUNK_MIN: *value = t.chunk_min; return GA_OK;
    default: GA_REQUIRE(false, GA_ERR_INVALID, "unknown tuning key %d", key);
  }
}

std::mutex g_default_mu;
}  // namespace
}  // namespace ga

extern "C" int ga_set_tuning(int32_t key, int32_t value) {
  clear_error();
  std::lock_guard<std::mutex> lk(g_default_mu);
  return tuning_set(tuning(), key, value);
}

extern "C" int ga_get_tuning(int32_t key, int32_t *value) {
  clear_error();
  std::lock_guard<std::mutex> lk(g_default_mu);
  return tuning_get(tuning(), key, value);
}

extern "C" int ga_plan_set_tuning(ga_plan *p, int32_t key, int32_t value) {
  clear_error();
  GA_REQUIRE(p != nullptr, GA_ERR_INVALID, "plan is NULL");
  return tuning_set(p->tune, key, value);
}

extern "C" int ga_plan_get_tuning(const ga_plan *p, int32_t key, int32_t *value) {
  clear_error();
  GA_REQUIRE(p != nullptr, GA_ERR_INVALID, "plan is NULL");
  return tuning_get(p->tune, key, value);
}

extern "C" int ga_plan_create(int32_t nside, int32_t lmax, ga_plan **out) {
  clear_error();
  GA_REQUIRE(out != nullptr, GA_ERR_INVALID, "plan output pointer is NULL");
  GA_REQUIRE(nside >= 1 && nside <= 8192, GA_ERR_UNSUPPORTED, "nside %d outside 1..8192", nside);
  GA_REQUIRE(lmax >= 0 && lmax <= 16384, GA_ERR_UNSUPPORTED, "lmax %d outside 0..16384", lmax);
  ga_plan *p = new (std::nothrow) ga_plan();
  GA_REQUIRE(p != nullptr, GA_ERR_NOMEM, "host allocation of plan failed");
  p->nside = nside;
  p->lmax = lmax;
  p->nrings = int32_t(nrings_of(nside));
  p->npairs = 2 * nside;
  p->npix = npix_of(nside);
  p->nalm = nalm_of(lmax);
  {
    std::lock_guard<std::mutex> lk(g_default_mu);
    p->tune = tuning();  // the process default at creation; per plan from here on
  }
  (void)hipGetDevice(&p->device);
  int rc = GA_OK;
  auto fail = [&](int code) { ga_plan_destroy(p); return code; };

  // ---- geometry
  std::vector<double> z(size_t(p->npairs)), sth(size_t(p->npairs));
  for (int rp = 0; rp < p->npairs; ++rp) {
    RingGeom g = ring_geom(nside, rp + 1);
    z[size_t(rp)] = g.z;
    sth[size_t(rp)] = g.sth;
  }
  std::vector<int64_t> start(size_t(p->nrings));
  std::vector<int32_t> rn(size_t(p->nrings)), pnum(size_t(p->nrings)), pden(size_t(p->nrings));
  for (int r = 0; r < p->nrings; ++r) {
    RingGeom g = ring_geom(nside, r + 1);
    start[size_t(r)] = g.start;
    rn[size_t(r)] = g.n;
    pnum[size_t(r)] = g.pnum;
    pden[size_t(r)] = g.pden;
  }
  if ((rc = upload(&p->d_z, z.data(), z.size() * 8, &p->table_bytes))) return fail(rc);
  if ((rc = upload(&p->d_sth, sth.data(), sth.size() * 8, &p->table_bytes))) return fail(rc);
  if ((rc = upload(&p->d_ring_start, start.data(), start.size() * 8, &p->table_bytes))) return fail(rc);
  if ((rc = upload(&p->d_ring_n, rn.data(), rn.size() * 4, &p->table_bytes))) return fail(rc);
  if ((rc = upload(&p->d_ring_pnum, pnum.data(), pnum.size() * 4, &p->table_bytes))) return fail(rc);
  if ((rc = upload(&p->d_ring_pden, pden.data(), pden.size() * 4, &p->table_bytes))) return fail(rc);

  // ---- Legendre tables
  {
    std::vector<double2> ab(size_t(p->nalm + 4), double2{0.0, 0.0});
    for (int64_t m = 0; m <= lmax; ++m) {
      const double mm = double(m);
      // (1, -1) at l = m: one step from (lambda_{m-2}, lambda_{m-1}) = (seed, 0)
      // yields lambda_mm, so kernels can step uniformly from l = m
      ab[size_t(hp_index(lmax, m, m))] = double2{1.0, -1.0};
      for (int64_t l = m + 1; l <= lmax; ++l) {
        const double ll = double(l);
        double a = std::sqrt((4.0 * ll * ll - 1.0) / (ll * ll - mm * mm));
        double b = std::sqrt(((ll - 1.0) * (ll - 1.0) - mm * mm) / (4.0 * (ll - 1.0) * (ll - 1.0) - 1.0));
        ab[size_t(hp_index(lmax, l, m))] = double2{a, b};
      }
    }
    if ((rc = upload(&p->d_ab, ab.data(), ab.size() * sizeof(double2), &p->table_bytes))) return fail(rc);
    // normalised form for the split-FP16 synthesis: y_l = lambda_l / c_l
    {
      // gam is padded by 64 entries: the synthesis loads a whole slab of 32
      // coefficients (and the next slab's) without clamping at the end of m;
      // entries past an m's last row belong to the next m (or the padding)
      // and only feed recurrence steps after that m's last row, whose values
      // are discarded
      std::vector<double> gam(size_t(p->nalm + 64), 0.0), csc(size_t(p->nalm + 4), 1.0);
      double bound = 0.0;
      for (int64_t m = 0; m <= lmax; ++m) {
        const int64_t b0 = hp_index(lmax, m, m);
        double cprev2 = 1.0, cprev = 1.0;  // c_{l-2}, c_{l-1}
        gam[size_t(b0)] = 0.0;
        csc[size_t(b0)] = 1.0;
        bound = std::max(bound, std::sqrt((2.0 * double(m) + 1.0) / (4.0 * M_PI)));
        for (int64_t l = m + 1; l <= lmax; ++l) {
          const double2 v = ab[size_t(b0 + (l - m))];
          const double c = l == m + 1 ? 1.0 : v.x * v.y * cprev2;
          gam[size_t(b0 + (l - m))] = v.x * cprev / c;
          csc[size_t(b0 + (l - m))] = c;
          bound = std::max(bound, std::sqrt((2.0 * double(l) + 1.0) / (4.0 * M_PI)) / c);
          cprev2 = cprev;
          cprev = c;
        }
      }
      int e = 8;
      while (e > 0 && std::ldexp(bound, e) >= 0x1p15 * M_SQRT2) --e;
      p->lam_exp16 = e;
      if ((rc = upload(&p->d_gam, gam.data(), gam.size() * 8, &p->table_bytes))) return fail(rc);
      if ((rc = upload(&p->d_csc, csc.data(), csc.size() * 8, &p->table_bytes))) return fail(rc);
    }
    // the same recurrence as (a, a b): lambda_l = (a x) lambda_{l-1} - (a b) lambda_{l-2},
    // one FMA on the dependency chain per step
    for (double2 &v : ab) v.y *= v.x;
    if ((rc = upload(&p->d_aab, ab.data(), ab.size() * sizeof(double2), &p->table_bytes))) return fail(rc);
    // c_m = (-1)^m sqrt((2m+1)/(4 pi) prod_{k<=m} (2k-1)/(2k)): lambda_mm = c_m sin^m
    std::vector<double> cm(size_t(lmax + 1));
    long double acc = 1.0L;
    for (int64_t m = 0; m <= lmax; ++m) {
      if (m > 0) acc *= (2.0L * m - 1.0L) / (2.0L * m);
      double v = double(std::sqrt((2.0L * m + 1.0L) / (4.0L * (long double)M_PI) * acc));
      cm[size_t(m)] = (m & 1) ? -v : v;
    }
    if ((rc = upload(&p->d_cm, cm.data(), cm.size() * 8, &p->table_bytes))) return fail(rc);
    std::vector<int64_t> so(size_t(lmax + 2));
    int64_t ns = 0;
    for (int64_t m = 0; m <= lmax; ++m) {
      so[size_t(m)] = ns;
      ns += nslab_of(lmax, m);
    }
    so[size_t(lmax + 1)] = ns;
    p->nslab = ns;
    if ((rc = upload(&p->d_slab_off, so.data(), so.size() * 8, &p->table_bytes))) return fail(rc);
    // where each wave of the split-FP16 synthesis starts its slab loop, and
    // the recurrence state there (computed on the device from the tables
    // above, by the same code the synthesis would run)
    p->f16_ngroups = (p->npairs + 63) / 64;
    const size_t nst = size_t(lmax + 1) * size_t(p->f16_ngroups) * 64;
    if ((rc = alloc_table(&p->d_f16_s0, size_t(lmax + 1) * size_t(p->f16_ngroups) * 4,
                          &p->table_bytes)) ||
        (rc = alloc_table(&p->d_f16_st, nst * sizeof(double2), &p->table_bytes)) ||
        (rc = alloc_table(&p->d_f16_k, nst * 4, &p->table_bytes)))
      return fail(rc);
    if ((rc = launch_legendre_f16_start(p, nullptr))) return fail(rc);
    const hipError_t se = hipStreamSynchronize(nullptr);
    if (se != hipSuccess) {
      set_error("building the start table failed: %s", hipGetErrorString(se));
      return fail(GA_ERR_HIP);
    }
  }

  // ---- ring FFT tables: every ring has n = 4 r pixels
  {
    int32_t max_r = 0;
    for (int r = 0; r < p->nrings; ++r) max_r = std::max(max_r, rn[size_t(r)] / 4);
    p->max_r = max_r;
    std::vector<int64_t> off(size_t(max_r + 1), -1);
    std::vector<char> used(size_t(max_r + 1), 0);
    for (int r = 0; r < p->nrings; ++r) used[size_t(rn[size_t(r)] / 4)] = 1;
    int64_t total = 0, maxL = 1, maxn_pow2 = 1;
    for (int32_t r = 1; r <= max_r; ++r) {
      if (!used[size_t(r)]) continue;
      if (is_pow2(r)) {
        maxL = std::max<int64_t>(maxL, r);
        maxn_pow2 = std::max<int64_t>(maxn_pow2, 4 * int64_t(r));
        continue;
      }
      int64_t L = blu_len(r);
      maxL = std::max(maxL, L);
      off[size_t(r)] = total;
      total += r + L;
    }
    p->max_L = int32_t(maxL);
    std::vector<double2> blu(size_t(std::max<int64_t>(total, 1)));
    for (int32_t r = 1; r <= max_r; ++r) {
      if (off[size_t(r)] < 0) continue;
      const int64_t L = blu_len(r);
      int bits = 0;
      while ((int64_t(1) << bits) < L) ++bits;
      double2 *c = blu.data() + off[size_t(r)];
      double2 *bh = c + r;
      std::vector<std::complex<double>> b(size_t(L), 0.0);
      for (int64_t k = 0; k < r; ++k) {
        int64_t q = (k * k) % (2 * r);
        double ang = M_PI * double(q) / double(r);
        c[k] = double2{std::cos(ang), std::sin(ang)};
        std::complex<double> cc(std::cos(ang), -std::sin(ang));
        b[size_t(k)] = cc;
        if (k > 0) b[size_t(L - k)] = cc;
      }
      host_fft(b);
      for (int64_t i = 0; i < L; ++i) {
        auto v = b[size_t(i)] / double(L);
        bh[bitrev(unsigned(i), bits)] = double2{v.real(), v.imag()};
      }
    }
    std::vector<float2> blu32(blu.size());
    for (size_t i = 0; i < blu.size(); ++i) blu32[i] = float2{float(blu[i].x), float(blu[i].y)};
    const int64_t twlen = std::max(maxL, maxn_pow2);
    p->tw_len = int32_t(twlen);
    std::vector<double2> tw(static_cast<size_t>(twlen));
    std::vector<float2> tw32(static_cast<size_t>(twlen));
    for (int64_t j = 0; j < twlen; ++j) {
      double ang = 2.0 * M_PI * double(j) / double(twlen);
      tw[size_t(j)] = double2{std::cos(ang), std::sin(ang)};
      tw32[size_t(j)] = float2{float(tw[size_t(j)].x), float(tw[size_t(j)].y)};
    }
    if ((rc = upload(&p->d_blu_off, off.data(), off.size() * 8, &p->table_bytes))) return fail(rc);
    std::vector<int32_t> lp, lb;
    for (int32_t ring = 0; ring < p->nrings; ++ring) {
      const int32_t r = rn[size_t(ring)] / 4;
      if (is_pow2(r)) {
        lp.push_back(ring);
        p->max_r_pow2 = std::max(p->max_r_pow2, r);
      } else {
        lb.push_back(ring);
        p->max_r_blu = std::max(p->max_r_blu, r);
        p->max_L_blu = std::max<int32_t>(p->max_L_blu, int32_t(blu_len(r)));
      }
    }
    p->n_pow2 = int32_t(lp.size());
    p->n_blu = int32_t(lb.size());
    std::vector<int32_t> lw, ln;
    const int32_t rw = p->max_r_pow2;
    const bool wide_ok = rw >= 256 && rw <= 4096 && 4 * int64_t(rw) > lmax;
    for (int32_t ring : lp) {
      const int32_t r = rn[size_t(ring)] / 4;
      if (wide_ok && r == rw) {
        lw.push_back(ring);
      } else {
        ln.push_back(ring);
        p->max_r_narrow = std::max(p->max_r_narrow, r);
      }
    }
    std::vector<int32_t> lbw, lbn;
    for (int32_t ring : lb) {
      const int32_t r = rn[size_t(ring)] / 4;
      const int32_t L = int32_t(blu_len(r));
      if (4 * int64_t(r) > lmax && r <= 2048 && L <= 4096) {
        lbw.push_back(ring);
        p->max_r_bwide = std::max(p->max_r_bwide, r);
        p->max_L_bwide = std::max(p->max_L_bwide, L);
      } else {
        lbn.push_back(ring);
        p->max_r_bnarrow = std::max(p->max_r_bnarrow, r);
        p->max_L_bnarrow = std::max(p->max_L_bnarrow, L);
      }
    }
    p->n_bwide = int32_t(lbw.size());
    p->n_bnarrow = int32_t(lbn.size());
    if (!lbw.empty() && (rc = upload(&p->d_rings_bwide, lbw.data(), lbw.size() * 4, &p->table_bytes)))
      return fail(rc);
    if (!lbn.empty() && (rc = upload(&p->d_rings_bnarrow, lbn.data(), lbn.size() * 4, &p->table_bytes)))
      return fail(rc);
    p->n_wide = int32_t(lw.size());
    p->n_narrow = int32_t(ln.size());
    p->r_wide = wide_ok ? rw : 0;
    if (!lw.empty() && (rc = upload(&p->d_rings_wide, lw.data(), lw.size() * 4, &p->table_bytes)))
      return fail(rc);
    if (!ln.empty() && (rc = upload(&p->d_rings_narrow, ln.data(), ln.size() * 4, &p->table_bytes)))
      return fail(rc);
    if (!lp.empty() && (rc = upload(&p->d_rings_pow2, lp.data(), lp.size() * 4, &p->table_bytes)))
      return fail(rc);
    if (!lb.empty() && (rc = upload(&p->d_rings_blu, lb.data(), lb.size() * 4, &p->table_bytes)))
      return fail(rc);
    if ((rc = upload(&p->d_blu64, blu.data(), blu.size() * sizeof(double2), &p->table_bytes))) return fail(rc);
    if ((rc = upload(&p->d_blu32, blu32.data(), blu32.size() * sizeof(float2), &p->table_bytes))) return fail(rc);
    if ((rc = upload(&p->d_tw64, tw.data(), tw.size() * sizeof(double2), &p->table_bytes))) return fail(rc);
    if ((rc = upload(&p->d_tw32, tw32.data(), tw32.size() * sizeof(float2), &p->table_bytes))) return fail(rc);
  }
  *out = p;
  return GA_OK;
}

extern "C" int ga_plan_destroy(ga_plan *p) {
  if (!p) return GA_OK;
  void *ptrs[] = {p->d_z, p->d_sth, p->d_ring_start, p->d_ring_n, p->d_ring_pnum, p->d_ring_pden,
                  p->d_ab, p->d_aab, p->d_gam, p->d_csc, p->d_cm, p->d_slab_off, p->d_spin[0], p->d_spin[1], p->d_spin[2],
                  p->d_spin_seed[0], p->d_spin_seed[1], p->d_spin_seed[2], p->d_blu_off,
                  p->d_blu64, p->d_blu32, p->d_tw64, p->d_tw32, p->d_rings_pow2,
                  p->d_rings_blu, p->d_rings_wide, p->d_rings_narrow, p->d_rings_bwide,
                  p->d_rings_bnarrow, p->d_f16_s0, p->d_f16_st, p->d_f16_k};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  delete p;
  return GA_OK;
}

extern "C" int ga_plan_table_bytes(const ga_plan *p, size_t *bytes) {
  GA_REQUIRE(p && bytes, GA_ERR_INVALID, "NULL argument");
  *bytes = p->table_bytes;
  return GA_OK;
}

extern "C" int ga_workspace_size(const ga_plan *p, int32_t nbatch, int32_t real_bytes,
                                 size_t *bytes) {
  GA_REQUIRE(p && bytes, GA_ERR_INVALID, "NULL argument");
  GA_REQUIRE(nbatch >= 1, GA_ERR_INVALID, "nbatch must be >= 1");
  GA_REQUIRE(real_bytes == 4 || real_bytes == 8, GA_ERR_INVALID, "real_bytes must be 4 or 8");
  if (real_bytes == 4) {
    *bytes = f32_workspace(p, nbatch).total;
    return GA_OK;
  }
  *bytes = size_t(p->nrings) * size_t(p->lmax + 1) * size_t(nbatch) * 2 * size_t(real_bytes);
  return GA_OK;
}

namespace ga {
f32_ws_layout f32_workspace(const ga_plan *p, int nbatch) {
  return f32_workspace(p, nbatch, p->tune);
}

f32_ws_layout f32_workspace(const ga_plan *p, int nbatch, const ga_tuning &tu) {
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  f32_ws_layout w;
  // the ring buffer holds whole groups of ring_group(nbatch) shells (and at
  // least a multiple of 8)
  const size_t gq = size_t(std::max(8, ring_group(nbatch)));
  const size_t nb = (size_t(nbatch) + gq - 1) / gq * gq;
  w.f_bytes = up(size_t(p->nrings) * size_t(p->lmax + 1) * nb * 8);
  const size_t nct = size_t(2 * nbatch + 31) / 32;  // 32-column tiles
  // per (slab, parity, tile): 64 lanes x (8 high + 8 low halves) = 2 KiB
  w.b_off = w.f_bytes;
  w.b_bytes = up(size_t(p->nslab) * 2 * nct * 2048);
  w.e_off = w.b_off + w.b_bytes;
  w.p_off = w.e_off + up(nct * 32 * 4 * 2);  // exponents + maxima scratch
  w.p_bytes = 0;
  const int acc_form = f16_accumulation(p, tu, nbatch);
  if (acc_form == -2) {
    // two chunks park the first in the ring buffer itself; only the O half
    // of the equator row needs a place of its own: (lmax + 1) x 2 nbatch
    w.p_bytes = up(size_t(p->lmax + 1) * size_t(2 * nbatch) * 4);
  } else if (acc_form == -3) {
    // two chunks park the first in a slot per (m, ring group, 64-column
    // group, wave) of the two-tile launch grid: 32 float4 per lane; the
    // one-tile grid needs at most as much.  Only the m whose l-range of
    // ceil((lmax - m + 1) / 32) slabs reaches the split threshold park:
    // m < lmax + 1 - 32 (chunk_min - 1)
    const int64_t cmin = tu.chunk_min > 0 ? tu.chunk_min : kTwoChunkMinSlabs;
    const int64_t mpark = std::max<int64_t>(0, int64_t(p->lmax) + 1 - 32 * (cmin - 1));
    const size_t nrg = size_t(p->npairs + 255) / 256;
    const size_t ncg = size_t(2 * nbatch + 63) / 64;
    w.p_bytes = up(size_t(mpark) * nrg * ncg * 4 * 64 * 32 * 16);
  }
  w.total = w.p_off + w.p_bytes;
  return w;
}

bool f16_park_fits(const ga_plan *p, int nbatch) {
  const uint64_t cg = uint64_t(2 * ring_group(nbatch));  // real columns per group
  const uint64_t rstride = uint64_t((2 * nbatch + cg - 1) / cg) * uint64_t(p->lmax + 1) * cg * 4;
  return 64 * rstride < (uint64_t(1) << 32);
}

int f16_accumulation(const ga_plan *p, const ga_tuning &tu, int nbatch) {
  (void)nbatch;
  if (tu.f16_flush != 0) return tu.f16_flush;
  return p->lmax > kAutoFlushLmax ? kAutoFlushSlabs : kAutoChunks;
}
}  // namespace ga
