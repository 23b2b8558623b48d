// Autotuner for the SHT launch shapes and the mixed-size batched request.
//
// The reference serves mixed HEALPix sizes by calling SHTCUDA once per size;
// its FFT plan cache is keyed on (n, nside) and is rebuilt whenever either
// changes (cuhpx/hpx_fft.cpp:49-58), so alternating sizes rebuild plans on
// every call.  Here each size keeps its own plan, and one request over all
// sizes is captured once into a HIP graph: every size is a branch of the graph
// (ring analysis -> Legendre analysis -> Legendre synthesis -> ring synthesis),
// the branches run concurrently, and a call is one hipGraphLaunch.  The launch
// shapes inside each branch are the ones hpx_sht_autotune measured as fastest
// on this device for that plan and batch.

#include "hpxmi_sht_internal.h"
#include "../../include/hpxmi.h"

#include <algorithm>
#include <vector>

namespace hpxmi {
namespace {

inline ShtPlan* W(hpx_sht_plan* h) { return reinterpret_cast<ShtPlan*>(h); }
inline const ShtPlan* R(const hpx_sht_plan* h) { return reinterpret_cast<const ShtPlan*>(h); }

constexpr int64_t kAlign = 256;
inline int64_t align_up(int64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Bytes written before each timed trial so that the trial starts with cold
// caches (twice the MI355X Infinity Cache), as each stage does inside a
// full transform, where the other stages have evicted its inputs.
constexpr size_t kFlushBytes = size_t(512) << 20;

// Time of one cold run of fn on s, best of three trials (ms).
template <class F>
int time_ms(F fn, hipStream_t s, hipEvent_t e0, hipEvent_t e1, void* flush, float* out) {
  float best = 1e30f;
  for (int t = 0; t < 3; ++t) {
    HPX_HIP(hipMemsetAsync(flush, t, kFlushBytes, s));
    HPX_HIP(hipEventRecord(e0, s));
    const int rc = fn();
    if (rc != HPX_OK) return rc;
    HPX_HIP(hipEventRecord(e1, s));
    HPX_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    HPX_HIP(hipEventElapsedTime(&ms, e0, e1));
    best = std::min(best, ms);
  }
  *out = best;
  return HPX_OK;
}

int validate(const ShtPlan* p, int precision, int64_t nbatch, int64_t ws_bytes) {
  if (p == nullptr) HPX_FAIL(HPX_ERR_ARG, "null plan");
  if (p->lmax < 1) HPX_FAIL(HPX_ERR_ARG, "plan was built with lmax = 0 (ring FFTs only)");
  if (precision < 0 || precision > 2) HPX_FAIL(HPX_ERR_ARG, "unknown precision %d", precision);
  if (nbatch < 0) HPX_FAIL(HPX_ERR_ARG, "nbatch must be >= 0");
  if (nbatch > max_maps(*p))
    HPX_FAIL(HPX_ERR_UNSUPPORTED, "%lld maps in one call exceed this plan's limit of %lld; split the batch",
             (long long)nbatch, (long long)max_maps(*p));
  const int64_t need = workspace_floats(*p, nbatch) * (int64_t)sizeof(float);
  if (ws_bytes < need)
    HPX_FAIL(HPX_ERR_ARG, "workspace of %lld bytes is smaller than the %lld bytes needed",
             (long long)ws_bytes, (long long)need);
  return HPX_OK;
}

struct Scratch {
  float* f = nullptr;
  float* alm = nullptr;
  void* flush = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~Scratch() {
    if (f) (void)hipFree(f);
    if (alm) (void)hipFree(alm);
    if (flush) (void)hipFree(flush);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

int autotune(ShtPlan& p, int precision, int64_t nbatch, float* ws, int64_t ws_bytes,
             hipStream_t s) {
  const int64_t ncol = col_pad(nbatch);
  Scratch t;
  HPX_HIP(hipMalloc(reinterpret_cast<void**>(&t.f), nbatch * p.npix * sizeof(float)));
  HPX_HIP(hipMalloc(reinterpret_cast<void**>(&t.alm), nbatch * p.lmax * (int64_t)p.mmax * 2 * sizeof(float)));
  HPX_HIP(hipMalloc(&t.flush, kFlushBytes));
  HPX_HIP(hipEventCreate(&t.e0));
  HPX_HIP(hipEventCreate(&t.e1));
  HPX_HIP(hipMemsetAsync(t.f, 0, nbatch * p.npix * sizeof(float), s));
  HPX_HIP(hipMemsetAsync(t.alm, 0, nbatch * p.lmax * (int64_t)p.mmax * 2 * sizeof(float), s));
  HPX_HIP(hipMemsetAsync(ws, 0, ws_bytes, s));
  float* X = ws;
  float* aT = ws + x_floats(p, nbatch);

  FftTune ft;
  for (int dir = 0; dir < 2; ++dir)
    for (int c = 0; c < 3; ++c) {
      if (p.class_n[c] == 0) continue;
      float best = 1e30f;
      for (int w : kFftWaveChoices[c])
        for (int tpf : kFftTpfChoices) {
          if (w * tpf > 1024) continue;
          float ms = 0.f;
          const int rc = time_ms(
              [&] {
                return dir == 0 ? launch_fft_class(p, true, c, w, tpf, t.f, X, nbatch, 1, 1.f, 0, s, 0)
                                : launch_fft_class(p, false, c, w, tpf, X, t.f, nbatch, 0, 1.f, 0, s, 0);
              },
              s, t.e0, t.e1, t.flush, &ms);
          if (rc == HPX_ERR_UNSUPPORTED) continue;  // shape does not fit this plan's LDS
          if (rc != HPX_OK) return rc;
          if (ms < best) {
            best = ms;
            ft.waves[dir][c] = w;
            ft.tpf[dir][c] = tpf;
          }
        }
    }
  GemmTune gt;
  float best_f = 1e30f, best_i = 1e30f;
  for (int ct : kGemmCtChoices) {
    float ms = 0.f;
    int rc = time_ms([&] { return launch_gemm_fwd(p, precision, ct, X, aT, nbatch, s); }, s, t.e0, t.e1, t.flush, &ms);
    if (rc != HPX_OK) return rc;
    if (ms < best_f) {
      best_f = ms;
      gt.fwd_ct = ct;
    }
    if (ct == kGemmCtChoices[1] && !p.vector && p.kpad % 64 == 0) {
      // the 8-wave forward kernel (128 columns per work-group)
      rc = time_ms([&] { return launch_gemm_fwd(p, precision, kGemmFwdW8, X, aT, nbatch, s); }, s, t.e0, t.e1,
                   t.flush, &ms);
      if (rc != HPX_OK) return rc;
      if (ms < best_f) {
        best_f = ms;
        gt.fwd_ct = kGemmFwdW8;
      }
    }
    for (int rh : kGemmRhChoices) {
      if (p.vector && rh != 1) continue;  // vector GEMMs run 64 ring pairs per work-group
      rc = time_ms([&] { return launch_gemm_inv(p, precision, ct, rh, aT, X, nbatch, s); }, s, t.e0, t.e1, t.flush,
                   &ms);
      if (rc != HPX_OK) return rc;
      if (ms < best_i) {
        best_i = ms;
        gt.inv_ct = ct;
        gt.inv_rh = rh;
      }
    }
  }
  HPX_HIP(hipStreamSynchronize(s));
  std::lock_guard<std::mutex> g(p.tune_mu);
  p.fft_tunes[ncol] = ft;
  p.gemm_tunes[{precision, ncol}] = gt;
  return HPX_OK;
}

// The launches of one size's request (the body of one graph branch).
int enqueue_request(const ShtPlan& p, int precision, int direction, const float* maps_in,
                    float* alm, float* maps_out, int64_t nbatch, float* ws, hipStream_t s) {
  int rc = HPX_OK;
  if (direction & HPX_MIXED_FORWARD) {
    if ((rc = launch_analysis(p, maps_in, ws, nbatch, 1, 1.0f, 0, s)) != HPX_OK) return rc;
    if ((rc = launch_legendre_fwd(p, precision, ws, alm, nbatch, s)) != HPX_OK) return rc;
  }
  if (direction & HPX_MIXED_INVERSE) {
    if ((rc = launch_legendre_inv(p, precision, alm, ws, nbatch, s)) != HPX_OK) return rc;
    if ((rc = launch_synthesis(p, ws, maps_out, nbatch, 0, 1.0f, 0, s)) != HPX_OK) return rc;
  }
  return HPX_OK;
}

struct ShtMixed {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int nbranch = 0;
};

// Streams and events used only while capturing.
struct CaptureResources {
  hipStream_t cs = nullptr;
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> ev;
  ~CaptureResources() {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
    for (hipStream_t x : side)
      if (x) (void)hipStreamDestroy(x);
    if (cs) (void)hipStreamDestroy(cs);
  }
};

}  // namespace
}  // namespace hpxmi

using namespace hpxmi;

extern "C" int hpx_sht_autotune(hpx_sht_plan* plan, int precision, int64_t nbatch, void* workspace,
                                int64_t workspace_bytes, hipStream_t stream) {
  ShtPlan* p = W(plan);
  int rc = validate(p, precision, nbatch, workspace_bytes);
  if (rc != HPX_OK || nbatch == 0) return rc;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  HPX_HIP(hipStreamIsCapturing(stream, &st));
  if (st != hipStreamCaptureStatusNone) HPX_FAIL(HPX_ERR_ARG, "autotune cannot run inside graph capture");
  return autotune(*p, precision, nbatch, static_cast<float*>(workspace), workspace_bytes, stream);
}

extern "C" int hpx_sht_tuned_shapes(const hpx_sht_plan* plan, int precision, int64_t nbatch,
                                    int32_t* shapes) {
  const ShtPlan* p = R(plan);
  if (p == nullptr || shapes == nullptr) HPX_FAIL(HPX_ERR_ARG, "null argument");
  if (precision < 0 || precision > 2) HPX_FAIL(HPX_ERR_ARG, "unknown precision %d", precision);
  const int64_t ncol = col_pad(nbatch);
  const FftTune f = p->fft_tune(ncol);
  const GemmTune g = p->gemm_tune(precision, ncol);
  for (int d = 0; d < 2; ++d)
    for (int c = 0; c < 3; ++c) {
      shapes[3 * d + c] = f.waves[d][c];
      shapes[9 + 3 * d + c] = f.tpf[d][c];
    }
  shapes[6] = g.fwd_ct;
  shapes[7] = g.inv_ct;
  shapes[8] = g.inv_rh;
  return HPX_OK;
}

extern "C" int hpx_sht_set_shapes(hpx_sht_plan* plan, int precision, int64_t nbatch, const int32_t* shapes) {
  ShtPlan* p = W(plan);
  if (p == nullptr || shapes == nullptr) HPX_FAIL(HPX_ERR_ARG, "null argument");
  if (precision < 0 || precision > 2) HPX_FAIL(HPX_ERR_ARG, "unknown precision %d", precision);
  if (nbatch < 1) HPX_FAIL(HPX_ERR_ARG, "nbatch must be >= 1");
  FftTune f;
  for (int d = 0; d < 2; ++d)
    for (int c = 0; c < 3; ++c) {
      const int w = shapes[3 * d + c], tpf = shapes[9 + 3 * d + c];
      if (w != kFftWaveChoices[c][0] && w != kFftWaveChoices[c][1] && w != kFftWaveChoices[c][2])
        HPX_FAIL(HPX_ERR_ARG, "no FFT kernel with %d fields per work-group for size class %d", w, c);
      if ((tpf != 64 && tpf != 128 && tpf != 256) || w * tpf > 1024)
        HPX_FAIL(HPX_ERR_ARG, "no FFT kernel with %d fields of %d threads", w, tpf);
      if (fft_lds_bytes(*p, c, w) > 160 * 1024)
        HPX_FAIL(HPX_ERR_UNSUPPORTED, "%d fields of size class %d exceed the LDS of a work-group", w, c);
      f.waves[d][c] = w;
      f.tpf[d][c] = tpf;
    }
  GemmTune g;
  g.fwd_ct = shapes[6];
  g.inv_ct = shapes[7];
  g.inv_rh = shapes[8];
  const bool fwd_ok = g.fwd_ct == kGemmCtChoices[0] || g.fwd_ct == kGemmCtChoices[1] ||
                      (g.fwd_ct == kGemmFwdW8 && !p->vector && p->kpad % 64 == 0);
  if (!fwd_ok) HPX_FAIL(HPX_ERR_ARG, "no forward GEMM with column tiles %d for this plan", g.fwd_ct);
  if (g.inv_ct != kGemmCtChoices[0] && g.inv_ct != kGemmCtChoices[1])
    HPX_FAIL(HPX_ERR_ARG, "no inverse GEMM with %d column tiles", g.inv_ct);
  if (g.inv_rh != kGemmRhChoices[0] && g.inv_rh != kGemmRhChoices[1])
    HPX_FAIL(HPX_ERR_ARG, "no inverse GEMM with %d row halves", g.inv_rh);
  const int64_t ncol = col_pad(nbatch);
  std::lock_guard<std::mutex> lk(p->tune_mu);
  p->fft_tunes[ncol] = f;
  p->gemm_tunes[{precision, ncol}] = g;
  return HPX_OK;
}

extern "C" int64_t hpx_sht_mixed_workspace_bytes(int nplans, const hpx_sht_plan* const* plans,
                                                 const int64_t* nbatch) {
  if (nplans < 0 || (nplans > 0 && (plans == nullptr || nbatch == nullptr))) return -1;
  int64_t total = 0;
  for (int i = 0; i < nplans; ++i) {
    const ShtPlan* p = R(plans[i]);
    if (p == nullptr || nbatch[i] < 0) return -1;
    total += align_up(workspace_floats(*p, nbatch[i]) * (int64_t)sizeof(float));
  }
  return total;
}

extern "C" int hpx_sht_mixed_create(int nplans, const hpx_sht_plan* const* plans,
                                    const int64_t* nbatch, int precision, int direction,
                                    const float* const* maps_in, float* const* alm,
                                    float* const* maps_out, void* workspace,
                                    int64_t workspace_bytes, hpx_sht_mixed** out) {
  if (out == nullptr) HPX_FAIL(HPX_ERR_ARG, "null output handle");
  *out = nullptr;
  if (nplans < 1 || nplans > HPX_MIXED_MAX_PLANS)
    HPX_FAIL(HPX_ERR_ARG, "nplans must be in [1, %d]", HPX_MIXED_MAX_PLANS);
  if (plans == nullptr || nbatch == nullptr || alm == nullptr) HPX_FAIL(HPX_ERR_ARG, "null argument");
  if (direction < 1 || direction > 3) HPX_FAIL(HPX_ERR_ARG, "direction must be 1, 2 or 3");
  if ((direction & HPX_MIXED_FORWARD) && maps_in == nullptr) HPX_FAIL(HPX_ERR_ARG, "null maps_in");
  if ((direction & HPX_MIXED_INVERSE) && maps_out == nullptr) HPX_FAIL(HPX_ERR_ARG, "null maps_out");
  const int64_t need = hpx_sht_mixed_workspace_bytes(nplans, plans, nbatch);
  if (need < 0) HPX_FAIL(HPX_ERR_ARG, "null plan or negative batch");
  if (workspace_bytes < need)
    HPX_FAIL(HPX_ERR_ARG, "workspace of %lld bytes is smaller than the %lld bytes needed",
             (long long)workspace_bytes, (long long)need);
  for (int i = 0; i < nplans; ++i) {
    int rc = validate(R(plans[i]), precision, nbatch[i], INT64_MAX);
    if (rc != HPX_OK) return rc;
    if (R(plans[i])->vector) HPX_FAIL(HPX_ERR_ARG, "plan %d is a vector plan; mixed requests take scalar plans", i);
  }

  CaptureResources cr;
  HPX_HIP(hipStreamCreateWithFlags(&cr.cs, hipStreamNonBlocking));
  cr.side.assign(nplans, nullptr);
  cr.ev.assign(nplans + 1, nullptr);
  for (int i = 0; i < nplans; ++i) HPX_HIP(hipStreamCreateWithFlags(&cr.side[i], hipStreamNonBlocking));
  for (int i = 0; i <= nplans; ++i) HPX_HIP(hipEventCreateWithFlags(&cr.ev[i], hipEventDisableTiming));

  HPX_HIP(hipStreamBeginCapture(cr.cs, hipStreamCaptureModeThreadLocal));
  int rc = HPX_OK;
  hipError_t e = hipEventRecord(cr.ev[nplans], cr.cs);
  char* ws = static_cast<char*>(workspace);
  for (int i = 0; i < nplans && rc == HPX_OK && e == hipSuccess; ++i) {
    const ShtPlan& p = *R(plans[i]);
    e = hipStreamWaitEvent(cr.side[i], cr.ev[nplans], 0);
    if (e != hipSuccess) break;
    if (nbatch[i] > 0)
      rc = enqueue_request(p, precision, direction, maps_in ? maps_in[i] : nullptr, alm[i],
                           maps_out ? maps_out[i] : nullptr, nbatch[i], reinterpret_cast<float*>(ws),
                           cr.side[i]);
    ws += align_up(workspace_floats(p, nbatch[i]) * (int64_t)sizeof(float));
    if (rc != HPX_OK) break;
    e = hipEventRecord(cr.ev[i], cr.side[i]);
    if (e == hipSuccess) e = hipStreamWaitEvent(cr.cs, cr.ev[i], 0);
  }
  hipGraph_t graph = nullptr;
  const hipError_t ee = hipStreamEndCapture(cr.cs, &graph);
  if (rc != HPX_OK || e != hipSuccess || ee != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    if (rc != HPX_OK) return rc;
    HPX_FAIL(HPX_ERR_HIP, "graph capture failed: %s",
             hipGetErrorString(e != hipSuccess ? e : ee));
  }
  ShtMixed* m = new ShtMixed();
  m->graph = graph;
  m->nbranch = nplans;
  e = hipGraphInstantiate(&m->exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(graph);
    delete m;
    HPX_FAIL(HPX_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
  }
  *out = reinterpret_cast<hpx_sht_mixed*>(m);
  return HPX_OK;
}

extern "C" int hpx_sht_mixed_launch(hpx_sht_mixed* mixed, hipStream_t stream) {
  ShtMixed* m = reinterpret_cast<ShtMixed*>(mixed);
  if (m == nullptr || m->exec == nullptr) HPX_FAIL(HPX_ERR_ARG, "null mixed request");
  HPX_HIP(hipGraphLaunch(m->exec, stream));
  return HPX_OK;
}

extern "C" void hpx_sht_mixed_destroy(hpx_sht_mixed* mixed) {
  ShtMixed* m = reinterpret_cast<ShtMixed*>(mixed);
  if (m == nullptr) return;
  if (m->exec) (void)hipGraphExecDestroy(m->exec);
  if (m->graph) (void)hipGraphDestroy(m->graph);
  delete m;
}
