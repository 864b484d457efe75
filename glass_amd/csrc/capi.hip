// extern "C" entry points of the spherical-harmonic transforms: argument
// checks, workspace carving and the kernel sequence of each transform.
#include <mutex>

#include "common.h"
#include "device.h"
#include "kernels.h"

namespace ga {
namespace {

template <typename T>
__global__ void residual_kernel(int64_t n, const T *__restrict__ a, T *__restrict__ b) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i] - b[i];
}

// dst = src * w * c per pixel (dst may be src): the per-pixel quadrature
// weights of a weighted analysis, c undoing the analysis's uniform 4 pi / npix
template <typename T>
__global__ void weight_pixels_kernel(int64_t n, int64_t npix, const T *__restrict__ src,
                                     const double *__restrict__ w, double c, T *dst) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = T(double(src[i]) * (w[i % npix] * c));
}

// same size as ga_workspace_size: fp32 ring buffers hold whole groups of
// ring_group(nbatch) shells
size_t ring_bytes(const ga_plan *p, int nbatch, int rb) {
  size_t b = 0;
  (void)ga_workspace_size(p, nbatch, rb, &b);
  return b;
}
size_t partial_bytes(const ga_plan *p, int nbatch, int rb) {
  return size_t(anal_partial_slots(p)) * size_t(p->nalm) * size_t(nbatch) * 2 * size_t(rb);
}

}  // namespace
}  // namespace ga

using namespace ga;

#define GA_PLAN_OK(p)                                                                      \
  do {                                                                                     \
    clear_error();                                                                         \
    GA_REQUIRE((p) != nullptr, GA_ERR_INVALID, "plan is NULL");                            \
  } while (0)

extern "C" int ga_map2alm_workspace_size(const ga_plan *p, int32_t nbatch, int32_t rb,
                                         size_t *bytes) {
  GA_PLAN_OK(p);
  GA_REQUIRE(bytes && nbatch >= 1 && (rb == 4 || rb == 8), GA_ERR_INVALID, "invalid argument");
  *bytes = ring_bytes(p, nbatch, rb) + partial_bytes(p, nbatch, rb);
  return GA_OK;
}

extern "C" int ga_plan_prepare_spin(ga_plan *p, int32_t spin) {
  GA_PLAN_OK(p);
  GA_REQUIRE(spin == 1 || spin == 2, GA_ERR_UNSUPPORTED, "spin %d not supported (1 or 2)", spin);
  return plan_spin_tables(p, spin);
}

extern "C" int ga_alm2map_f64(const ga_plan *p, const double *alm, double *map, void *ws,
                              size_t ws_bytes, void *stream) {
  GA_PLAN_OK(p);
  GA_REQUIRE(alm && map && ws, GA_ERR_INVALID, "NULL buffer");
  GA_REQUIRE(ws_bytes >= ring_bytes(p, 1, 8), GA_ERR_WORKSPACE,
             "workspace %zu < %zu bytes", ws_bytes, ring_bytes(p, 1, 8));
  hipStream_t s = (hipStream_t)stream;
  double *F = (double *)ws;
  int rc = launch_legendre_synth_f64(p, alm, F, s);
  if (rc) return rc;
  return launch_ring_synth_f64(p, 1, 1, F, map, nullptr, s);
}

extern "C" int ga_alm2map_batch_f32(const ga_plan *p, int32_t S, const float *alm, float *maps,
                                    const ga_epilogue *epi, void *ws, size_t ws_bytes,
                                    void *stream) {
  GA_PLAN_OK(p);
  GA_REQUIRE(S >= 1 && alm && maps && ws, GA_ERR_INVALID, "invalid argument");
  GA_REQUIRE(ws_bytes >= ring_bytes(p, S, 4), GA_ERR_WORKSPACE,
             "workspace %zu < %zu bytes", ws_bytes, ring_bytes(p, S, 4));
  hipStream_t s = (hipStream_t)stream;
  int rc = launch_legendre_synth_batch(p, S, alm, ws, s);
  if (rc) return rc;
  return launch_ring_synth_f32(p, S, ring_group(S), (const float *)ws, maps, epi, s);
}

// ---- fixed device scratch of ga_alm2map_batch_f32_noalloc -------------------
// One array per device (each GPU that loads the library holds its own copy
// until the process ends).  Host code never passes the array's name to a
// kernel: its address on the current device comes from hipGetSymbolAddress,
// taken once per device on the entry point's first call there.
namespace ga {
namespace {
constexpr size_t kNoAllocScratchBytes = size_t(128) << 20;
constexpr int kMaxDevices = 64;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

int scratch_address(int dev, void **out) {
  GA_REQUIRE(dev >= 0 && dev < kMaxDevices, GA_ERR_UNSUPPORTED, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    GA_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    GA_REQUIRE(ptr != nullptr, GA_ERR_HIP, "no device address for the scratch array");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return GA_OK;
}
}  // namespace
}  // namespace ga

extern "C" int ga_noalloc_scratch_bytes(size_t *bytes) {
  clear_error();
  GA_REQUIRE(bytes, GA_ERR_INVALID, "NULL argument");
  *bytes = kNoAllocScratchBytes;
  return GA_OK;
}

extern "C" int ga_alm2map_batch_f32_noalloc(const ga_plan *p, int32_t S, const float *alm,
                                            float *maps, const ga_epilogue *epi, void *stream) {
  GA_PLAN_OK(p);
  GA_REQUIRE(S >= 1 && alm && maps, GA_ERR_INVALID, "invalid argument");
  // the compact two-chunk form (first chunk parked in the ring buffer) in
  // place of the one with a region of its own: bit-identical results, and the
  // scratch holds twice the batch
  ga_tuning tu = p->tune;
  if (f16_accumulation(p, tu, S) == -3) tu.f16_flush = -2;
  const size_t need = f32_workspace(p, S, tu).total;
  GA_REQUIRE(need <= kNoAllocScratchBytes, GA_ERR_UNSUPPORTED,
             "nside %d, lmax %d, %d maps need %zu bytes of scratch; the fixed scratch holds %zu",
             p->nside, p->lmax, S, need, kNoAllocScratchBytes);
  int dev = -1;
  GA_CHECK_HIP(hipGetDevice(&dev));
  GA_REQUIRE(dev == p->device, GA_ERR_INVALID, "plan belongs to device %d, current device is %d",
             p->device, dev);
  void *ws = nullptr;
  int rc = scratch_address(dev, &ws);
  if (rc) return rc;
  // the kernels of ga_alm2map_batch_f32 on the fixed scratch: everything they
  // need zeroed (the column maxima) is cleared by a fill kernel in the
  // preparation launch, every other byte they read was written earlier in
  // the same call
  hipStream_t s = (hipStream_t)stream;
  rc = launch_legendre_synth_batch(p, tu, S, alm, ws, s);
  if (rc) return rc;
  return launch_ring_synth_f32(p, S, ring_group(S), (const float *)ws, maps, epi, s);
}

extern "C" int ga_legendre_synth_batch_f32(const ga_plan *p, int32_t S, const float *alm,
                                           void *ws, size_t ws_bytes, void *stream) {
  GA_PLAN_OK(p);
  GA_REQUIRE(S >= 1 && alm && ws, GA_ERR_INVALID, "invalid argument");
  GA_REQUIRE(ws_bytes >= ring_bytes(p, S, 4), GA_ERR_WORKSPACE, "workspace %zu < %zu bytes",
             ws_bytes, ring_bytes(p, S, 4));
  return launch_legendre_synth_batch(p, S, alm, ws, (hipStream_t)stream);
}

extern "C" int ga_legendre_prepare_batch_f32(const ga_plan *p, int32_t S, const float *alm,
                                             void *ws, size_t ws_bytes, void *stream) {
  GA_PLAN_OK(p);
  GA_REQUIRE(S >= 1 && alm && ws, GA_ERR_INVALID, "invalid argument");
  GA_REQUIRE(ws_bytes >= ring_bytes(p, S, 4), GA_ERR_WORKSPACE, "workspace %zu < %zu bytes",
             ws_bytes, ring_bytes(p, S, 4));
  return launch_legendre_prepare_batch(p, p->tune, S, alm, ws, (hipStream_t)stream);
}

extern "C" int ga_legendre_prepare_colmax_batch_f32(const ga_plan *p, int32_t S,
                                                    const float *alm, const float *colpart,
                                                    int64_t nparts, void *ws, size_t ws_bytes,
                                                    void *stream) {
  GA_PLAN_OK(p);
  GA_REQUIRE(S >= 1 && alm && colpart && nparts >= 1 && ws, GA_ERR_INVALID, "invalid argument");
  GA_REQUIRE(ws_bytes >= ring_bytes(p, S, 4), GA_ERR_WORKSPACE, "workspace %zu < %zu bytes",
             ws_bytes, ring_bytes(p, S, 4));
  return launch_legendre_prepare_batch(p, p->tune, S, alm, ws, (hipStream_t)stream, colpart, nparts);
}

extern "C" int ga_legendre_contract_batch_f32(const ga_plan *p, int32_t S, const float *alm,
                                              void *ws, size_t ws_bytes, void *stream) {
  GA_PLAN_OK(p);
  GA_REQUIRE(S >= 1 && alm && ws, GA_ERR_INVALID, "invalid argument");
  GA_REQUIRE(ws_bytes >= ring_bytes(p, S, 4), GA_ERR_WORKSPACE, "workspace %zu < %zu bytes",
             ws_bytes, ring_bytes(p, S, 4));
  return launch_legendre_contract_batch(p, p->tune, S, alm, ws, (hipStream_t)stream);
}

extern "C" int ga_legendre_contract_grouped_f32(int32_t ngroup, const ga_plan *const *plans,
                                                const int32_t *nshell, void *const *workspace,
                                                const size_t *ws_bytes, void *stream) {
  clear_error();
  GA_REQUIRE(ngroup >= 1 && ngroup <= kGroupedMax && plans && nshell && workspace && ws_bytes,
             GA_ERR_INVALID, "invalid argument");
  for (int i = 0; i < ngroup; ++i) {
    GA_PLAN_OK(plans[i]);
    GA_REQUIRE(nshell[i] >= 1 && workspace[i], GA_ERR_INVALID, "invalid argument (geometry %d)",
               i);
    const size_t need = f32_workspace(plans[i], nshell[i]).total;
    GA_REQUIRE(ws_bytes[i] >= need, GA_ERR_WORKSPACE, "workspace %d: %zu < %zu bytes", i,
               ws_bytes[i], need);
  }
  return launch_legendre_contract_grouped(ngroup, plans, nshell, workspace, (hipStream_t)stream);
}

extern "C" int ga_ring_synth_batch_f32(const ga_plan *p, int32_t S, const float *F, float *maps,
                                       const ga_epilogue *epi, void *stream) {
  GA_PLAN_OK(p);
  GA_REQUIRE(S >= 1 && F && maps, GA_ERR_INVALID, "invalid argument");
  return launch_ring_synth_f32(p, S, ring_group(S), F, maps, epi, (hipStream_t)stream);
}

// pixw: per-pixel quadrature weights (npix doubles, summing to 4 pi) in place
// of the uniform 4 pi / npix, or NULL
template <typename T>
static int map2alm_impl(const ga_plan *p, int S, const T *map, T *alm, int niter, T *map_tmp,
                        void *ws, size_t ws_bytes, hipStream_t s,
                        const double *pixw = nullptr) {
  const int rb = int(sizeof(T));
  GA_REQUIRE(map && alm && ws && niter >= 0 && (niter == 0 || map_tmp), GA_ERR_INVALID,
             "invalid argument");
  GA_REQUIRE(!pixw || map_tmp, GA_ERR_INVALID, "a weighted analysis needs map_tmp");
  const double wc = double(p->npix) / (4.0 * M_PI);
  const size_t need = ring_bytes(p, S, rb) + partial_bytes(p, S, rb);
  GA_REQUIRE(ws_bytes >= need, GA_ERR_WORKSPACE, "workspace %zu < %zu bytes", ws_bytes, need);
  T *F = (T *)ws;
  T *partial = (T *)((char *)ws + ring_bytes(p, S, rb));
  const int64_t npx = p->npix * S;
  int rc;
  for (int it = 0; it <= niter; ++it) {
    const T *src = map;
    if (it > 0) {
      // map_tmp = map - S(alm)
      bool subtract = true;
      if (sizeof(T) == 8) {
        rc = launch_legendre_synth_f64(p, (const double *)alm, (double *)F, s);
        if (rc) return rc;
        rc = launch_ring_synth_f64(p, 1, 1, (const double *)F, (double *)map_tmp, nullptr, s);
      } else {
        rc = launch_legendre_synth_batch(p, S, (const float *)alm, ws, s);
        if (rc) return rc;
        // the subtraction fused into the ring stage's store where it can be
        rc = launch_ring_resid_f32(p, S, ring_group(S), (const float *)F, (float *)map_tmp,
                                   (const float *)map, s);
        if (rc == GA_OK) {
          subtract = false;
        } else if (rc == GA_ERR_UNSUPPORTED) {
          rc = launch_ring_synth_f32(p, S, ring_group(S), (const float *)F, (float *)map_tmp,
                                     nullptr, s);
        }
      }
      if (rc) return rc;
      if (subtract) {
        hipLaunchKernelGGL(residual_kernel<T>, dim3(unsigned((npx + 255) / 256)), dim3(256), 0,
                           s, npx, map, map_tmp);
        GA_LAUNCH_CHECK();
      }
      src = map_tmp;
    }
    if (pixw) {
      hipLaunchKernelGGL(weight_pixels_kernel<T>, dim3(unsigned((npx + 255) / 256)), dim3(256), 0,
                         s, npx, int64_t(p->npix), src, pixw, wc, map_tmp);
      GA_LAUNCH_CHECK();
      src = map_tmp;
    }
    if (sizeof(T) == 8) {
      rc = launch_ring_anal_f64(p, S, (const double *)src, (double *)F, s);
      if (rc) return rc;
      rc = launch_legendre_anal_f64(p, S, (const double *)F, (double *)alm, (double *)partial,
                                    it > 0, s);
    } else {
      rc = launch_ring_anal_f32(p, S, (const float *)src, (float *)F, s);
      if (rc) return rc;
      rc = launch_legendre_anal_f32(p, S, (const float *)F, (float *)alm, (float *)partial, it > 0,
                                    s);
    }
    if (rc) return rc;
  }
  return GA_OK;
}

extern "C" int ga_map2alm_f64(const ga_plan *p, const double *map, double *alm, int32_t niter,
                              double *map_tmp, void *ws, size_t ws_bytes, void *stream) {
  GA_PLAN_OK(p);
  return map2alm_impl<double>(p, 1, map, alm, niter, map_tmp, ws, ws_bytes, (hipStream_t)stream);
}

extern "C" int ga_map2alm_weighted_f64(const ga_plan *p, const double *map,
                                       const double *weights, double *alm, int32_t niter,
                                       double *map_tmp, void *ws, size_t ws_bytes, void *stream) {
  GA_PLAN_OK(p);
  GA_REQUIRE(weights, GA_ERR_INVALID, "NULL weights");
  return map2alm_impl<double>(p, 1, map, alm, niter, map_tmp, ws, ws_bytes, (hipStream_t)stream,
                              weights);
}

extern "C" int ga_map2alm_weighted_batch_f32(const ga_plan *p, int32_t S, const float *maps,
                                             const double *weights, float *alm, int32_t niter,
                                             float *map_tmp, void *ws, size_t ws_bytes,
                                             void *stream) {
  GA_PLAN_OK(p);
  GA_REQUIRE(S >= 1 && weights, GA_ERR_INVALID, "invalid argument");
  return map2alm_impl<float>(p, S, maps, alm, niter, map_tmp, ws, ws_bytes, (hipStream_t)stream,
                             weights);
}

extern "C" int ga_map2alm_batch_f32(const ga_plan *p, int32_t S, const float *maps, float *alm,
                                    int32_t niter, float *map_tmp, void *ws, size_t ws_bytes,
                                    void *stream) {
  GA_PLAN_OK(p);
  GA_REQUIRE(S >= 1, GA_ERR_INVALID, "nshell must be >= 1");
  return map2alm_impl<float>(p, S, maps, alm, niter, map_tmp, ws, ws_bytes, (hipStream_t)stream);
}

extern "C" int ga_alm2map_spin_f64(const ga_plan *p, int32_t spin, const double *E,
                                   const double *B, double *maps, void *ws, size_t ws_bytes,
                                   void *stream) {
  GA_PLAN_OK(p);
  GA_REQUIRE(E && B && maps && ws, GA_ERR_INVALID, "NULL buffer");
  GA_REQUIRE(ws_bytes >= ring_bytes(p, 2, 8), GA_ERR_WORKSPACE, "workspace %zu < %zu bytes",
             ws_bytes, ring_bytes(p, 2, 8));
  hipStream_t s = (hipStream_t)stream;
  int rc = launch_spin_synth_f64(p, spin, 1, E, B, (double *)ws, s);
  if (rc) return rc;
  return launch_ring_synth_f64(p, 2, 2, (const double *)ws, maps, nullptr, s);
}

extern "C" int ga_alm2map_spin_batch_f32(const ga_plan *p, int32_t spin, int32_t S,
                                         const float *E, const float *B, float *maps, void *ws,
                                         size_t ws_bytes, void *stream) {
  GA_PLAN_OK(p);
  GA_REQUIRE(S >= 1 && E && maps && ws, GA_ERR_INVALID, "invalid argument");
  GA_REQUIRE(ws_bytes >= ring_bytes(p, 2 * S, 4), GA_ERR_WORKSPACE, "workspace %zu < %zu bytes",
             ws_bytes, ring_bytes(p, 2 * S, 4));
  GA_REQUIRE(spin == 1 || spin == 2, GA_ERR_UNSUPPORTED, "spin %d not supported (1 or 2)", spin);
  hipStream_t s = (hipStream_t)stream;
  if (p->tune.spin_path == 1) {
    GA_REQUIRE(B, GA_ERR_INVALID, "the VALU spin path needs almB");
    int rc = launch_spin_synth_f32(p, spin, S, E, B, (float *)ws, s);
    if (rc) return rc;
    return launch_ring_synth_f32(p, 2 * S, 2 * S, (const float *)ws, maps, nullptr, s);
  }
  int rc = launch_spin_synth_f16(p, spin, S, E, B, ws, s);
  if (rc) return rc;
  return launch_ring_synth_f32(p, 2 * S, ring_group(2 * S), (const float *)ws, maps, nullptr, s);
}

// ---- graph capture ------------------------------------------------------------

struct ga_graph {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

extern "C" int ga_graph_begin(void *stream) {
  clear_error();
  GA_REQUIRE(stream, GA_ERR_INVALID, "capture needs a created stream, not the null stream");
  GA_CHECK_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
  return GA_OK;
}

extern "C" int ga_graph_end(void *stream, ga_graph **out) {
  clear_error();
  GA_REQUIRE(stream && out, GA_ERR_INVALID, "invalid argument");
  *out = nullptr;
  hipGraph_t g = nullptr;
  GA_CHECK_HIP(hipStreamEndCapture((hipStream_t)stream, &g));
  GA_REQUIRE(g, GA_ERR_HIP, "the capture recorded no graph");
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
    return GA_ERR_HIP;
  }
  *out = new ga_graph{g, exec};
  return GA_OK;
}

extern "C" int ga_graph_launch(const ga_graph *graph, void *stream) {
  clear_error();
  GA_REQUIRE(graph, GA_ERR_INVALID, "invalid argument");
  GA_CHECK_HIP(hipGraphLaunch(graph->exec, (hipStream_t)stream));
  return GA_OK;
}

extern "C" int ga_graph_destroy(ga_graph *graph) {
  clear_error();
  if (!graph) return GA_OK;
  const hipError_t e1 = hipGraphExecDestroy(graph->exec);
  const hipError_t e2 = hipGraphDestroy(graph->graph);
  delete graph;
  GA_REQUIRE(e1 == hipSuccess && e2 == hipSuccess, GA_ERR_HIP, "graph destroy failed");
  return GA_OK;
}
