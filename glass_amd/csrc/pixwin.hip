// HEALPix pixel window functions, computed from their definition on the GPU.
//
// The window of pixel p for spin s is the harmonic power of its normalised
// top hat, W^s_l(p) = (4 pi / (2l+1)) sum_m |<sY_lm>_p|^2, and the map window
// is w^s_l = sqrt(mean_p W^s_l(p)).  By the addition theorem of spin-weighted
// harmonics, (4 pi / (2l+1)) sum_m sY_lm(a) sY_lm(b)^* = d^l_ss(beta)
// e^{i s psi}, where beta is the angle between a and b and psi the rotation
// between the local (e_theta, e_phi) frame at b and the frame at a carried to
// b along the geodesic, so
//     W^s_l(p) = << d^l_ss(beta_ab) cos(s psi_ab) >>_{a, b in p}
// (d^l_00 = P_l(cos beta)).  Each pixel is integrated as two triangles (split
// along its centre line x + y = const of the face coordinates, where the
// HEALPix projection changes formula between the polar caps and the
// equatorial belt) in collapsed coordinates toward each apex; the apex of a
// polar pixel is the pole, where these coordinates are the smooth polar ones,
// so the quadrature converges spectrally (6 x 6 Gauss-Legendre nodes per
// triangle: 72 nodes, 2628 unordered pairs per pixel).  The face coordinates
// (x, y) are equal-area, so uniform weights in them are area weights.
//
// Symmetry: the 4 north faces are rotations of each other and the south faces
// mirror them (z -> -z), and within a face the pixels mirror across x = y;
// beta and cos(s psi) are invariant under all three.  Pixels of the
// equatorial belt (rings nside + 1 .. 3 nside - 1) are congruent within a
// ring.  So the pixels computed are those of face 0 with ix <= iy and
// ix + iy >= nside - 1 (rings 1 .. nside; multiplicity 8 or 16) plus one pixel
// per belt ring nside + 1 .. 2 nside (multiplicity 8 nside, 4 nside on the
// equator): about nside^2 / 4 pixels instead of 12 nside^2.
//
// Pairs are binned by beta (nb bins on [0, 2.4 / nside]) into LDS histograms
// of the weight and its first two moments in the bin, for w and w cos(s psi),
// as 64-bit fixed point (2^-fb units, fb = 61 - ceil(log2(12 nside^2)): at
// least 2^-31), so the sums do not depend on the order
// in which threads and workgroups add them.  Per bin, a three-node rule at the
// bin's edges and centre that matches the three moments replaces the pairs;
// the Legendre sums over the 2 nb + 1 nodes give W_l exactly up to the third
// moment of the within-bin distribution.
#include "common.h"
#include "device.h"
#include "kernels.h"

namespace ga {
namespace {

constexpr int kPwG = 6;                            // Gauss nodes per direction of a triangle
constexpr int kPwNodes = 2 * kPwG * kPwG;          // 72 nodes per pixel
constexpr int kPwPairs = kPwNodes * (kPwNodes + 1) / 2;  // 2628 unordered pairs
constexpr int kPwBins = 2048;
constexpr int kPwThreads = 512;
constexpr double kPwBetaMax = 2.4;                  // times 1 / nside: bound on a pixel's diameter

struct gl_rule {
  double u[kPwG], wu[kPwG];  // nodes / weights on [0, 1]
  double v[kPwG], wv[kPwG];  // nodes / weights on [-1, 1]
};

// HEALPix base-face offsets (Gorski et al. 2005): ring coordinate of the
// face's north vertex in units of nside (jr = jrll - x - y) and its phi offset
// in units of pi / 4
__constant__ int kJrll[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};
__constant__ int kJpll[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};

struct pw_node {
  double n[3], et[3], ep[2], w;  // unit vector, e_theta, e_phi (z part 0), weight
};

// point of face f at continuous face coordinates (x, y) in [0, 1]^2
__device__ void face_point(int f, double x, double y, pw_node &o) {
  const double jr = double(kJrll[f]) - x - y;
  double z, st, nr;
  if (jr < 1.0) {  // north polar cap: 1 - z = jr^2 / 3
    const double t = jr * jr / 3.0;
    z = 1.0 - t;
    st = sqrt(t * (2.0 - t));
    nr = jr;
  } else if (jr > 3.0) {  // south polar cap
    nr = 4.0 - jr;
    const double t = nr * nr / 3.0;
    z = t - 1.0;
    st = sqrt(t * (2.0 - t));
  } else {  // equatorial belt: z linear in jr
    z = (2.0 - jr) * (2.0 / 3.0);
    st = sqrt((1.0 - z) * (1.0 + z));
    nr = 1.0;
  }
  const double phi_q = nr > 0.0 ? (double(kJpll[f]) * nr + x - y) / nr : 0.0;  // phi / (pi/4)
  double sp, cp;
  sincospi(0.25 * phi_q, &sp, &cp);
  o.n[0] = st * cp;
  o.n[1] = st * sp;
  o.n[2] = z;
  o.et[0] = z * cp;
  o.et[1] = z * sp;
  o.et[2] = -st;
  o.ep[0] = -sp;
  o.ep[1] = cp;
}

// work item -> (face, ix, iy, multiplicity)
__device__ void pw_item(int64_t it, int nside, int64_t ntri, int &f, int &ix, int &iy,
                        double &mult) {
  if (it < ntri) {  // face 0, ix <= iy: it = iy (iy + 1) / 2 + ix
    int64_t q = int64_t(sqrt(2.0 * double(it) + 0.25) - 0.5);
    while (q * (q + 1) / 2 > it) --q;
    while ((q + 1) * (q + 2) / 2 <= it) ++q;
    f = 0;
    iy = int(q);
    ix = int(it - q * (q + 1) / 2);
    mult = ix == iy ? 8.0 : 16.0;
    if (ix + iy < nside - 1) mult = 0.0;  // belt pixel: counted by its ring below
  } else {  // belt ring i = nside + 1 + j, represented by face 4, ix = nside - 1
    const int j = int(it - ntri);
    const int ring = nside + 1 + j;
    f = 4;
    ix = nside - 1;
    iy = 2 * nside - ring;
    mult = ring == 2 * nside ? 4.0 * nside : 8.0 * nside;
  }
}

__device__ __forceinline__ long long to_fix(double v, double fix) {
  return __double2ll_rn(v * fix);
}

__global__ __launch_bounds__(kPwThreads, 1) void pixwin_pairs_kernel(
    int nside, int spin, int64_t ntri, int64_t nitems, gl_rule g, double dbeta_inv, double fix,
    unsigned long long *__restrict__ hist) {
  // hist layout: [6][kPwBins]: w, w d, w d^2, c, c d, c d^2 (c = w cos(s psi),
  // d = offset from the bin centre in bin widths)
  __shared__ unsigned long long h[6][kPwBins];
  __shared__ pw_node nd[kPwNodes];
  __shared__ uint16_t pairs[kPwPairs];
  const int tid = threadIdx.x;
  for (int i = tid; i < 6 * kPwBins; i += kPwThreads) (&h[0][0])[i] = 0ull;
  for (int a = tid; a < kPwNodes; a += kPwThreads) {
    // pairs (a, b >= a) of row a start at a kPwNodes - a (a - 1) / 2
    const int start = a * kPwNodes - a * (a - 1) / 2;
    for (int b = a; b < kPwNodes; ++b) pairs[start + b - a] = uint16_t(a * 128 + b);
  }
  const double h0 = 1.0 / double(nside);
  for (int64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
    int f, ix, iy;
    double mult;
    pw_item(it, nside, ntri, f, ix, iy, mult);
    if (mult == 0.0) continue;  // block-uniform
    __syncthreads();  // previous item's nodes are no longer read
    if (tid < kPwNodes) {
      const int tri = tid / (kPwG * kPwG), r = tid - tri * kPwG * kPwG;
      const int iu = r / kPwG, iv = r - iu * kPwG;
      const double sc = double(ix + iy + 1) * h0, dc = double(ix - iy) * h0;
      const double u = g.u[iu], hw = h0 * (1.0 - u);
      const double s = tri == 0 ? sc + h0 * u : sc - h0 * u;
      const double d = dc + hw * g.v[iv];
      pw_node o;
      face_point(f, 0.5 * (s + d), 0.5 * (s - d), o);
      // area weights (the triangle's width in d is 2 hw): they sum to 1 over
      // the pixel
      o.w = g.wu[iu] * (1.0 - u) * g.wv[iv] * 0.5;
      nd[tid] = o;
    }
    __syncthreads();
    for (int p = tid; p < kPwPairs; p += kPwThreads) {
      const int ab = pairs[p], a = ab >> 7, b = ab & 127;
      const pw_node &A = nd[a];
      const pw_node &B = nd[b];
      const double D0 = B.n[0] - A.n[0], D1 = B.n[1] - A.n[1], D2 = B.n[2] - A.n[2];
      const double ch2 = D0 * D0 + D1 * D1 + D2 * D2;  // chord^2 = 2 (1 - cos beta)
      const double beta = 2.0 * asin(fmin(1.0, 0.5 * sqrt(ch2)));
      double c2 = 1.0;
      if (spin != 0 && a != b) {
        // tangent at a toward b: D + q A, at b away from a: D - q B (q = 1 - cos beta)
        const double q = 0.5 * ch2;
        const double t1x = D0 + q * A.n[0], t1y = D1 + q * A.n[1], t1z = D2 + q * A.n[2];
        const double t2x = D0 - q * B.n[0], t2y = D1 - q * B.n[1], t2z = D2 - q * B.n[2];
        const double c1 = t1x * A.et[0] + t1y * A.et[1] + t1z * A.et[2];
        const double s1 = t1x * A.ep[0] + t1y * A.ep[1];
        const double cb = t2x * B.et[0] + t2y * B.et[1] + t2z * B.et[2];
        const double sb = t2x * B.ep[0] + t2y * B.ep[1];
        const double re = cb * c1 + sb * s1, im = sb * c1 - cb * s1;  // |.|^2 e^{i psi}
        const double nn = re * re + im * im;
        if (nn > 0.0) {
          // cos(s psi) for s = 1, 2
          c2 = spin == 2 ? (re * re - im * im) / nn : re / sqrt(nn);
        }
      }
      const double t = beta * dbeta_inv;
      const int j = min(int(t), kPwBins - 1);
      const double dd = t - (double(j) + 0.5);
      const double w = (a == b ? 1.0 : 2.0) * A.w * B.w * mult;
      const double wc = w * c2;
      atomicAdd(&h[0][j], (unsigned long long)to_fix(w, fix));
      atomicAdd(&h[1][j], (unsigned long long)to_fix(w * dd, fix));
      atomicAdd(&h[2][j], (unsigned long long)to_fix(w * dd * dd, fix));
      if (spin != 0) {
        atomicAdd(&h[3][j], (unsigned long long)to_fix(wc, fix));
        atomicAdd(&h[4][j], (unsigned long long)to_fix(wc * dd, fix));
        atomicAdd(&h[5][j], (unsigned long long)to_fix(wc * dd * dd, fix));
      }
    }
  }
  __syncthreads();
  for (int i = tid; i < 6 * kPwBins; i += kPwThreads) {
    const unsigned long long v = (&h[0][0])[i];
    if (v) atomicAdd(&hist[i], v);
  }
}

__global__ void pixwin_clear_kernel(unsigned long long *__restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// per node k = 0 .. 2 nb (beta_k = k / 2 bin widths): the weights of the
// three-node rule (nodes at the bin's left edge, centre, right edge) that
// matches the bin's moments m0, m1, m2 (offsets d in bin widths):
// left = 2 m2 - m1, centre = m0 - 4 m2, right = 2 m2 + m1; normalised by the
// exact integer total weight (so that w_0 = 1 whatever the fixed-point
// rounding of the pair weights)
__global__ void pixwin_nodes_kernel(const long long *__restrict__ hist,
                                    double *__restrict__ wn) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > 2 * kPwBins) return;
  long long total = 0;
  for (int j = 0; j < kPwBins; ++j) total += hist[j];
  const double norm = 1.0 / double(total);
  for (int q = 0; q < 2; ++q) {
    const long long *hq = hist + 3 * q * kPwBins;
    double v = 0.0;
    if (k & 1) {
      const int j = k >> 1;
      v = (double(hq[j]) - 4.0 * double(hq[2 * kPwBins + j])) * norm;
    } else {
      const int j = k >> 1;
      if (j > 0)  // right edge of bin j - 1
        v += (2.0 * double(hq[2 * kPwBins + j - 1]) + double(hq[kPwBins + j - 1])) * norm;
      if (j < kPwBins)  // left edge of bin j
        v += (2.0 * double(hq[2 * kPwBins + j]) - double(hq[kPwBins + j])) * norm;
    }
    wn[q * (2 * kPwBins + 1) + k] = v;
  }
}

// Legendre sums over the nodes: thread = node, recurrence over l; each wave
// reduces its 64 nodes per l (xor butterfly: every lane gets the same sum)
// into part[q][wave][l].
__global__ __launch_bounds__(256) void pixwin_sum_kernel(int lmax, int spin, double dbeta,
                                                         const double *__restrict__ wn,
                                                         double *__restrict__ part, int nwaves) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int wave = k >> 6, lane = k & 63;
  if (wave >= nwaves) return;  // whole waves past the last partial row
  const int nn = 2 * kPwBins + 1;
  const bool ok = k < nn;
  const double beta = 0.5 * double(min(k, nn - 1)) * dbeta;
  const double x = cos(beta);
  const double w0 = ok ? wn[k] : 0.0;
  const double w2 = ok ? wn[nn + k] : 0.0;
  // spin 0: P_l(x); spin s: d^l_ss(beta) by its three-term recurrence in l
  //   l ((l+1)^2 - s^2) d_{l+1} = (2l+1) (l (l+1) x - s^2) d_l - (l+1) (l^2 - s^2) d_{l-1}
  double p0 = 1.0, p1 = x;  // P_{l-1}, P_l
  const double s2 = double(spin) * double(spin);
  double q0 = 0.0, q1 = spin == 2 ? 0.25 * (1.0 + x) * (1.0 + x) : 0.5 * (1.0 + x);  // d_{s-1}, d_s
  double *__restrict__ out0 = part + int64_t(wave) * (lmax + 1);
  double *__restrict__ out2 = part + int64_t(nwaves + wave) * (lmax + 1);
  for (int l = 0; l <= lmax; ++l) {
    double P;
    if (l == 0) P = 1.0;
    else if (l == 1) P = x;
    else {
      const double pn = (double(2 * l - 1) * x * p1 - double(l - 1) * p0) / double(l);
      p0 = p1;
      p1 = pn;
      P = pn;
    }
    double d = 0.0;
    if (spin != 0 && l >= spin) {
      if (l > spin) {
        const double lm = double(l - 1);
        const double dn = (double(2 * l - 1) * (lm * (lm + 1.0) * x - s2) * q1 -
                           double(l) * (lm * lm - s2) * q0) /
                          (lm * (double(l) * double(l) - s2));
        q0 = q1;
        q1 = dn;
      }
      d = q1;
    }
    double a = w0 * P, c = w2 * d;
    for (int o = 32; o >= 1; o >>= 1) {
      a += __shfl_xor(a, o, 64);
      c += __shfl_xor(c, o, 64);
    }
    if (lane == 0) {
      out0[l] = a;
      out2[l] = c;
    }
  }
}

__global__ void pixwin_final_kernel(int lmax, int spin, int nwaves, const double *__restrict__ part,
                                    double *__restrict__ pw0, double *__restrict__ pws) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l > lmax) return;
  double a = 0.0, c = 0.0;
  for (int w = 0; w < nwaves; ++w) {
    a += part[int64_t(w) * (lmax + 1) + l];
    c += part[int64_t(nwaves + w) * (lmax + 1) + l];
  }
  pw0[l] = sqrt(fmax(a, 0.0));
  if (pws) pws[l] = l >= spin ? sqrt(fmax(c, 0.0)) : 0.0;
}

// Gauss-Legendre nodes on [-1, 1] by Newton's method on P_n
void gauss_legendre(int n, double *x, double *w) {
  for (int i = 0; i < n; ++i) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      const double dp = n * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (fabs(dz) < 1e-16) {
        p0 = 1.0;
        p1 = z;
        for (int k = 2; k <= n; ++k) {
          const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
          p0 = p1;
          p1 = p2;
        }
        const double dp2 = n * (z * p1 - p0) / (z * z - 1.0);
        x[n - 1 - i] = z;
        w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp2 * dp2);
        break;
      }
    }
  }
}

struct pw_ws {
  size_t hist_off, node_off, part_off, total;
  int nwaves;
};
pw_ws pixwin_layout(int lmax) {
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  pw_ws w;
  w.nwaves = (2 * kPwBins + 1 + 63) / 64;
  w.hist_off = 0;
  w.node_off = up(sizeof(long long) * 6 * kPwBins);
  w.part_off = w.node_off + up(sizeof(double) * 2 * (2 * kPwBins + 1));
  w.total = w.part_off + up(sizeof(double) * 2 * size_t(w.nwaves) * size_t(lmax + 1));
  return w;
}

constexpr int kPwMaxNside = 8192;  // keeps 2^-31 fixed-point resolution

}  // namespace
}  // namespace ga

using namespace ga;

extern "C" int ga_pixwin_workspace_size(int32_t nside, int32_t lmax, size_t *bytes) {
  clear_error();
  GA_REQUIRE(bytes && nside >= 1 && lmax >= 0, GA_ERR_INVALID, "invalid argument");
  *bytes = pixwin_layout(lmax).total;
  return GA_OK;
}

extern "C" int ga_pixwin(int32_t nside, int32_t lmax, int32_t spin, double *pw0, double *pws,
                         void *ws, size_t ws_bytes, void *stream) {
  clear_error();
  GA_REQUIRE(nside >= 1 && lmax >= 0 && pw0 && ws, GA_ERR_INVALID, "invalid argument");
  GA_REQUIRE(spin == 0 || spin == 1 || spin == 2, GA_ERR_UNSUPPORTED,
             "spin %d not supported (0, 1 or 2)", spin);
  GA_REQUIRE(spin == 0 || pws, GA_ERR_INVALID, "the spin window needs an output");
  GA_REQUIRE(nside <= kPwMaxNside, GA_ERR_UNSUPPORTED, "nside %d above %d", nside, kPwMaxNside);
  const pw_ws lay = pixwin_layout(lmax);
  GA_REQUIRE(ws_bytes >= lay.total, GA_ERR_WORKSPACE, "workspace %zu < %zu bytes", ws_bytes,
             lay.total);
  hipStream_t s = (hipStream_t)stream;
  char *base = static_cast<char *>(ws);
  auto *hist = reinterpret_cast<unsigned long long *>(base + lay.hist_off);
  auto *wn = reinterpret_cast<double *>(base + lay.node_off);
  auto *part = reinterpret_cast<double *>(base + lay.part_off);

  gl_rule g;
  double x[kPwG], w[kPwG];
  gauss_legendre(kPwG, x, w);
  for (int i = 0; i < kPwG; ++i) {
    g.v[i] = x[i];
    g.wv[i] = w[i];
    g.u[i] = 0.5 * (x[i] + 1.0);
    g.wu[i] = 0.5 * w[i];
  }
  // fixed point: 2^fb units per unit weight, fb as large as keeps the total
  // (12 nside^2) and every partial sum below 2^62
  const int fb = 61 - int(ceil(log2(12.0 * double(nside) * double(nside))));
  const double fix = ldexp(1.0, fb);
  const int64_t ntri = int64_t(nside) * (nside + 1) / 2;
  const int64_t nitems = ntri + nside;
  const double beta_max = kPwBetaMax / double(nside);
  const double dbeta = beta_max / double(kPwBins);

  hipLaunchKernelGGL(pixwin_clear_kernel, dim3((6 * kPwBins + 255) / 256), dim3(256), 0, s, hist,
                     6 * kPwBins);
  GA_LAUNCH_CHECK();
  const int grid = int(std::min<int64_t>(nitems, 2048));
  hipLaunchKernelGGL(pixwin_pairs_kernel, dim3(grid), dim3(kPwThreads), 0, s, nside, spin, ntri,
                     nitems, g, 1.0 / dbeta, fix, hist);
  GA_LAUNCH_CHECK();
  hipLaunchKernelGGL(pixwin_nodes_kernel, dim3((2 * kPwBins + 1 + 255) / 256), dim3(256), 0, s,
                     reinterpret_cast<const long long *>(hist), wn);
  GA_LAUNCH_CHECK();
  hipLaunchKernelGGL(pixwin_sum_kernel, dim3((lay.nwaves * 64 + 255) / 256), dim3(256), 0, s,
                     lmax, spin, dbeta, wn, part, lay.nwaves);
  GA_LAUNCH_CHECK();
  hipLaunchKernelGGL(pixwin_final_kernel, dim3((lmax + 1 + 255) / 256), dim3(256), 0, s, lmax, spin,
                     lay.nwaves, part, pw0, spin ? pws : nullptr);
  GA_LAUNCH_CHECK();
  return GA_OK;
}
