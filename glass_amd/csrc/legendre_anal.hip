// Legendre stage of the analysis (map2alm), VALU version, FP64 and FP32.
//
//   a_lm[s] = sum_rp lambda_lm(theta_rp) ( F_N[rp, m, s] + (-1)^(l+m) F_S[rp, m, s] )
//
// The contraction runs over ring pairs, which live in lanes, so each lane
// carries RPL ring pairs (their recurrences in FP64 and their even/odd
// combinations Q = F_N +- F_S in registers), forms its partial sum over them
// and the wave reduces the partial across lanes once per l.  Each wave writes
// its own partial [wave][idx][s]; a second kernel sums the waves in a fixed
// order, so the result does not depend on scheduling.
//
// The spin-s synthesis (alm2map_spin) stage lives here too: one lane per
// (m, ring pair) running the +s and -s Wigner-d recurrences, with the E/B alm
// read as wave-uniform values.
#include "common.h"
#include "device.h"
#include "kernels.h"

namespace ga {

namespace {

// Sum over the 64 lanes of a wave, complete in lane 63, with DPP lane moves
// (no LDS crossbar): quads, half rows, rows, then the row broadcasts 15 / 31.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_add(float v) {
  const int m = __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWMASK, 0xF, false);
  return v + __int_as_float(m);
}
__device__ __forceinline__ float wave_sum_last(float v) {
  v = dpp_add<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
  v = dpp_add<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
  v = dpp_add<0x141, 0xF>(v);  // row_half_mirror
  v = dpp_add<0x140, 0xF>(v);  // row_mirror: every lane holds its row's sum
  v = dpp_add<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v = dpp_add<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  return v;                    // lane 63: the wave's sum
}
__device__ __forceinline__ double wave_sum_last(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;  // every lane (the FP64 path keeps the exact butterfly order)
}

constexpr int kAnalRPL = 8;  // ring pairs per lane

template <typename T, int SC>
__global__ __launch_bounds__(256) void legendre_anal_kernel(
    int lmax, int npairs, int nrings, int S, const double *__restrict__ z,
    const double *__restrict__ sth, const double2 *__restrict__ ab, const double *__restrict__ cm,
    const cplx<T> *__restrict__ F, cplx<T> *__restrict__ partial, int64_t nalm) {
  constexpr int RPL = kAnalRPL;
  const int m = blockIdx.x;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int gw = blockIdx.y * 4 + wave;  // global wave index = partial slot
  if (gw * 64 * RPL >= npairs) return;   // wave-uniform; the kernel has no barriers
  const int s0 = blockIdx.z * SC;
  const int64_t base = hp_index(lmax, m, m);
  const double2 *__restrict__ co = ab + base;

  double x[RPL], p0[RPL], p1[RPL];
  int kk[RPL];
  cplx<T> qe[RPL][SC], qo[RPL][SC];
#pragma unroll
  for (int j = 0; j < RPL; ++j) {
    const int rp = gw * 64 * RPL + j * 64 + lane;
    const bool ok = rp < npairs;
    const int rpc = ok ? rp : npairs - 1;
    x[j] = z[rpc];
    p1[j] = 0.0;
    legendre_seed(cm[m], sth[rpc], m, p0[j], kk[j]);
#pragma unroll
    for (int c = 0; c < SC; ++c) {
      cplx<T> fn = {T(0), T(0)}, fs = {T(0), T(0)};
      if (ok && s0 + c < S) {
        fn = F[(int64_t(rp) * (lmax + 1) + m) * S + s0 + c];
        if (rp < npairs - 1) fs = F[(int64_t(nrings - 1 - rp) * (lmax + 1) + m) * S + s0 + c];
      }
      qe[j][c] = cadd(fn, fs);
      qo[j][c] = csub(fn, fs);
    }
  }
  cplx<T> *out = partial + (int64_t(gw) * nalm + base) * S + s0;
  for (int l = m; l <= lmax; ++l) {
    const double2 c = co[l - m];  // (1, -1) at l = m turns (seed, 0) into lambda_mm
    const bool odd = (l - m) & 1;
    T acc_re[SC], acc_im[SC];
#pragma unroll
    for (int cc = 0; cc < SC; ++cc) { acc_re[cc] = T(0); acc_im[cc] = T(0); }
    bool any_on = false;
#pragma unroll
    for (int j = 0; j < RPL; ++j) {
      const double pn = c.x * (x[j] * p1[j] - c.y * p0[j]);
      p0[j] = p1[j];
      p1[j] = pn;
      if (kk[j] < 0 && fabs(p1[j]) > 0x1p300) { p1[j] *= 0x1p-600; p0[j] *= 0x1p-600; ++kk[j]; }
      const T v = kk[j] == 0 ? T(p1[j]) : T(0);
      any_on |= kk[j] == 0;
#pragma unroll
      for (int cc = 0; cc < SC; ++cc) {
        const cplx<T> q = odd ? qo[j][cc] : qe[j][cc];
        acc_re[cc] += v * q.re;
        acc_im[cc] += v * q.im;
      }
    }
    if (wave_any(any_on)) {
#pragma unroll
      for (int cc = 0; cc < SC; ++cc) {
        acc_re[cc] = wave_sum_last(acc_re[cc]);
        acc_im[cc] = wave_sum_last(acc_im[cc]);
      }
    }
    if (lane == 63) {
#pragma unroll
      for (int cc = 0; cc < SC; ++cc)
        if (s0 + cc < S) out[int64_t(l - m) * S + cc] = cplx<T>{acc_re[cc], m == 0 ? T(0) : acc_im[cc]};
    }
  }
}

// alm[idx][s] (+)= sum_w partial[w][idx][s], fixed order
template <typename T>
__global__ void reduce_partials_kernel(int64_t n, int nw, const cplx<T> *__restrict__ partial,
                                       cplx<T> *__restrict__ alm, int accumulate) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cplx<T> s = {T(0), T(0)};
  for (int w = 0; w < nw; ++w) s = cadd(s, partial[int64_t(w) * n + i]);
  if (accumulate) s = cadd(s, alm[i]);
  alm[i] = s;
}

// ---------------------------------------------------------------------------
// spin-s synthesis: Q_m = sum_l E F+ + i B F-,  U_m = sum_l B F+ - i E F-,
// F+- = -(sY +- (-1)^s (-s)Y) / 2; south ring via F+(pi - t) = sigma F+(t),
// F-(pi - t) = -sigma F-(t), sigma = (-1)^(l+m+s).

template <typename T, int SC>
__global__ __launch_bounds__(256) void spin_synth_kernel(
    int lmax, int npairs, int nrings, int spin, int S, const double *__restrict__ z,
    const double *__restrict__ sth, const double *__restrict__ tab /* (A,B,C) per (l,m) */,
    const double2 *__restrict__ seedk /* per m: (+s mant, +s exp), */,
    const double2 *__restrict__ seedkm /* per m: (-s mant, -s exp) */,
    const int32_t *__restrict__ ring_pnum, const int32_t *__restrict__ ring_pden,
    const cplx<T> *__restrict__ E, const cplx<T> *__restrict__ B, cplx<T> *__restrict__ Fo) {
  const int m = blockIdx.x;
  const int rp = blockIdx.y * blockDim.x + threadIdx.x;
  const int s0 = blockIdx.z * SC;
  const int rpc = rp < npairs ? rp : npairs - 1;
  const double x = z[rpc], st = sth[rpc];
  const int l0 = m > spin ? m : spin;
  // accumulators [shell][QA, QB, UA, UB] x [parity class of l + m + s]
  cplx<T> acc[SC][4][2];
#pragma unroll
  for (int c = 0; c < SC; ++c)
#pragma unroll
    for (int a = 0; a < 4; ++a) acc[c][a][0] = acc[c][a][1] = cplx<T>{T(0), T(0)};

  if (l0 <= lmax) {
    double sh = sqrt(0.5 * (1.0 - x)), ch = sqrt(0.5 * (1.0 + x));
    if (x > 0.5) sh = st / (2.0 * ch);
    else if (x < -0.5) ch = st / (2.0 * sh);
    const int dif = m > spin ? m - spin : spin - m;
    double ms1, ms2, mc1, mc2;
    long long es1, es2, ec1, ec2;
    scaled_power(sh, m + spin, ms1, es1);
    scaled_power(ch, dif, mc1, ec1);
    scaled_power(sh, dif, ms2, es2);
    scaled_power(ch, m + spin, mc2, ec2);
    const double2 kp = seedk[m], km = seedkm[m];
    double pp, pm;
    int kp_, km_;
    {
      int t;
      double v = frexp(kp.x * ms1 * mc1, &t);
      to_scaled(v, (long long)kp.y + es1 + ec1 + t, pp, kp_);
      v = frexp(km.x * ms2 * mc2, &t);
      to_scaled(v, (long long)km.y + es2 + ec2 + t, pm, km_);
    }
    double a0 = pp, a1 = 0.0, b0 = pm, b1 = 0.0;  // (lambda_{l-2}, lambda_{l-1}) for +s, -s
    const double sgn = (spin & 1) ? -1.0 : 1.0;
    const int64_t base = hp_index(lmax, m, m);
    for (int l = l0; l <= lmax; ++l) {
      const double *cf = tab + 3 * (base + (l - m));
      const double A = cf[0], AB = cf[1], C = cf[2];
      const double n1 = fma(fma(A, x, AB), a1, -(C * a0));
      const double n2 = fma(fma(A, x, -AB), b1, -(C * b0));
      a0 = a1; a1 = n1;
      b0 = b1; b1 = n2;
      if (kp_ < 0 && fabs(a1) > 0x1p300) { a1 *= 0x1p-600; a0 *= 0x1p-600; ++kp_; }
      if (km_ < 0 && fabs(b1) > 0x1p300) { b1 *= 0x1p-600; b0 *= 0x1p-600; ++km_; }
      const double lp = kp_ == 0 ? a1 : 0.0, lm = km_ == 0 ? b1 : 0.0;
      const T fp = T(-0.5 * (lp + sgn * lm)), fm = T(-0.5 * (lp - sgn * lm));
      const int par = (l + m + spin) & 1;
      const int64_t ia = (base + (l - m)) * S + s0;
#pragma unroll
      for (int c = 0; c < SC; ++c) {
        if (s0 + c >= S) break;
        cplx<T> e = E[ia + c], b = B[ia + c];
        if (m == 0) { e.im = T(0); b.im = T(0); }
        // QA += E F+, QB += i B F-, UA += B F+, UB += -i E F-
        acc[c][0][par] = cadd(acc[c][0][par], cplx<T>{e.re * fp, e.im * fp});
        acc[c][1][par] = cadd(acc[c][1][par], cplx<T>{-b.im * fm, b.re * fm});
        acc[c][2][par] = cadd(acc[c][2][par], cplx<T>{b.re * fp, b.im * fp});
        acc[c][3][par] = cadd(acc[c][3][par], cplx<T>{e.im * fm, -e.re * fm});
      }
    }
  }
  if (rp >= npairs) return;
  // ring phase e^{i m phi0} applied here (the same on the south mirror)
  const int pnum = ring_pnum[rp], pden = ring_pden[rp];
  double ps = 0.0, pc = 1.0;
  if (pnum) sincospi(2.0 * double((int64_t(m) * pnum) % pden) / double(pden), &ps, &pc);
  const cplx<T> ph = {T(pc), T(ps)};
  // output ring buffer [ring][m][2S]: column 2s holds Q_s, 2s + 1 holds U_s
  const int64_t rowN = (int64_t(rp) * (lmax + 1) + m) * (2 * S) + 2 * s0;
  const int64_t rowS = (int64_t(nrings - 1 - rp) * (lmax + 1) + m) * (2 * S) + 2 * s0;
#pragma unroll
  for (int c = 0; c < SC; ++c) {
    if (s0 + c >= S) break;
    const cplx<T> qa_e = acc[c][0][0], qb_e = acc[c][1][0], qa_o = acc[c][0][1], qb_o = acc[c][1][1];
    const cplx<T> ua_e = acc[c][2][0], ub_e = acc[c][3][0], ua_o = acc[c][2][1], ub_o = acc[c][3][1];
    Fo[rowN + 2 * c] = cmul(cadd(cadd(qa_e, qb_e), cadd(qa_o, qb_o)), ph);
    Fo[rowN + 2 * c + 1] = cmul(cadd(cadd(ua_e, ub_e), cadd(ua_o, ub_o)), ph);
    if (rp < npairs - 1) {
      Fo[rowS + 2 * c] = cmul(cadd(csub(qa_e, qb_e), csub(qb_o, qa_o)), ph);
      Fo[rowS + 2 * c + 1] = cmul(cadd(csub(ua_e, ub_e), csub(ub_o, ua_o)), ph);
    }
  }
}

}  // namespace

int anal_partial_slots(const ga_plan *p) { return (p->npairs + 64 * kAnalRPL - 1) / (64 * kAnalRPL); }

template <typename T>
static int launch_anal(const ga_plan *p, int S, const T *F, T *alm, T *partial, bool accumulate,
                       hipStream_t stream) {
  const int nw = anal_partial_slots(p);
  if (sizeof(T) == 4 && p->tune.anal_path != 1) {
    int rc = launch_legendre_anal_f16(p, S, (const float *)F, (float *)partial, stream);
    if (rc) return rc;
  } else {
    const int nblk_y = (nw + 3) / 4;
    constexpr int SC = sizeof(T) == 8 ? 1 : 2;
    dim3 grid(unsigned(p->lmax + 1), unsigned(nblk_y), unsigned((S + SC - 1) / SC));
    hipLaunchKernelGGL((legendre_anal_kernel<T, SC>), grid, dim3(256), 0, stream, p->lmax,
                       p->npairs, p->nrings, S, p->d_z, p->d_sth, p->d_ab, p->d_cm,
                       (const cplx<T> *)F, (cplx<T> *)partial, p->nalm);
    GA_LAUNCH_CHECK();
  }
  const int64_t n = p->nalm * S;
  hipLaunchKernelGGL(reduce_partials_kernel<T>, dim3(unsigned((n + 255) / 256)), dim3(256), 0,
                     stream, n, nw, (const cplx<T> *)partial, (cplx<T> *)alm, accumulate ? 1 : 0);
  GA_LAUNCH_CHECK();
  return GA_OK;
}

int launch_legendre_anal_f64(const ga_plan *p, int S, const double *F, double *alm,
                             double *partial, bool accumulate, hipStream_t s) {
  return launch_anal<double>(p, S, F, alm, partial, accumulate, s);
}
int launch_legendre_anal_f32(const ga_plan *p, int S, const float *F, float *alm, float *partial,
                             bool accumulate, hipStream_t s) {
  return launch_anal<float>(p, S, F, alm, partial, accumulate, s);
}

template <typename T>
static int launch_spin(const ga_plan *p, int spin, int S, const T *E, const T *B, T *Fo,
                       hipStream_t stream) {
  GA_REQUIRE(spin >= 1 && spin <= 2 && p->d_spin[spin] && p->d_spin_seed[spin], GA_ERR_INVALID,
             "spin %d tables not prepared (call ga_plan_prepare_spin)", spin);
  constexpr int SC = sizeof(T) == 8 ? 1 : 4;
  dim3 grid(unsigned(p->lmax + 1), unsigned((p->npairs + 255) / 256), unsigned((S + SC - 1) / SC));
  hipLaunchKernelGGL((spin_synth_kernel<T, SC>), grid, dim3(256), 0, stream, p->lmax, p->npairs,
                     p->nrings, spin, S, p->d_z, p->d_sth, p->d_spin[spin],
                     p->d_spin_seed[spin], p->d_spin_seed[spin] + (p->lmax + 1), p->d_ring_pnum,
                     p->d_ring_pden, (const cplx<T> *)E, (const cplx<T> *)B, (cplx<T> *)Fo);
  GA_LAUNCH_CHECK();
  return GA_OK;
}

int launch_spin_synth_f64(const ga_plan *p, int spin, int S, const double *E, const double *B,
                          double *Fo, hipStream_t s) {
  return launch_spin<double>(p, spin, S, E, B, Fo, s);
}
int launch_spin_synth_f32(const ga_plan *p, int spin, int S, const float *E, const float *B,
                          float *Fo, hipStream_t s) {
  return launch_spin<float>(p, spin, S, E, B, Fo, s);
}

}  // namespace ga
