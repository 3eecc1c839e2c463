// C-ABI for the contractions and the dielectric inversion (include/pygw_hip.h).
#include "common.h"

#include <map>
#include <mutex>
#include <tuple>

int chi0_f64_launch(const void* M, long strideK, int np, int pb, const double* kcw, int nom, int k0, int nk,
                    double coef, void* eps, int beta, int nw, hipStream_t stream);
int chi0_mixed_launch(const void* M, long strideK, int np, int pb, const double* kcw, int nom, int k0, int nkc,
                      double coef, int G, int beta, void* ws, size_t ws_bytes, void* eps, hipStream_t stream);

// Autotune cache: (device, precision, pb, n_pairs, n_om) -> schedule s in
// {1, 2, 4}: FP64, s frequencies per 64x64 tile block; split precision, the
// FP32 tile accumulator folded into FP64 every 64*s chunks of 32 pairs.  The
// only global state of the library; guarded by a mutex.
namespace {
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int, int>, int> g_tune;

int tuned_nw(int precision, int pb, int np, int nom) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(std::make_tuple(dev, precision, pb, np, nom));
  return it == g_tune.end() ? 2 : it->second;
}
}  // namespace
size_t chi0_mixed_ws_bytes(int pb, int np, int nom, int nk);
struct Chi0BatchPlan;
size_t chi0_batch_ws_bytes(int nq, const int* pb, const int* np, int nom);
int chi0_batch_plan(int nq, const void* const* M, const int* pb, const int* np, const double* const* kcw, int nom,
                    double coef, void* const* eps, void* ws, size_t ws_bytes, hipStream_t stream,
                    Chi0BatchPlan** out);
int chi0_batch_run(const Chi0BatchPlan* P, int G, hipStream_t stream);
void chi0_batch_free(Chi0BatchPlan* P);
int eps_inverse_launch(void* eps, int n, int nom, void* ws, int* dinfo, hipStream_t stream);
size_t eps_inverse_ws_bytes(int n, int nom);

extern "C" int pygw_chi0_workspace(int pb, int n_pairs, int n_om, int n_k, int precision, size_t* bytes) {
  if (pb <= 0 || n_pairs <= 0 || n_om <= 0 || n_k <= 0 || !bytes || (precision != 0 && precision != 1))
    return PYGW_E_ARG;
  *bytes = precision == 0 ? 0 : chi0_mixed_ws_bytes(pb, n_pairs, n_om, n_k);
  return PYGW_OK;
}

extern "C" int pygw_chi0_accumulate(const void* M, long stride_k, int n_pairs, int pb, const double* kcw, int n_om,
                                    int k0, int n_k, double coef, int precision, int beta, void* eps,
                                    void* workspace, size_t ws_bytes, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (!M || !kcw || !eps || n_pairs <= 0 || pb <= 0 || n_om <= 0 || n_k <= 0 || k0 < 0) return PYGW_E_ARG;
  if (stride_k < (long)n_pairs * pb || (beta != 0 && beta != 1)) return PYGW_E_ARG;
  const int nw = tuned_nw(precision, pb, n_pairs, n_om);
  if (precision == 0)
    return chi0_f64_launch(M, stride_k, n_pairs, pb, kcw, n_om, k0, n_k, coef, eps, beta, nw, stream);
  if (precision != 1) return PYGW_E_ARG;
  void* ws = workspace;
  bool own = false;
  if (!ws) {
    ws_bytes = chi0_mixed_ws_bytes(pb, n_pairs, n_om, n_k);
    HIP_TRY(hipMallocAsync(&ws, ws_bytes, stream));
    own = true;
  }
  // a smaller caller workspace walks k in smaller slabs (rejected below one k)
  int rc = chi0_mixed_launch(M, stride_k, n_pairs, pb, kcw, n_om, k0, n_k, coef, 64 * nw, beta, ws, ws_bytes, eps,
                             stream);
  if (own) {
    hipError_t e = hipFreeAsync(ws, stream);
    if (e != hipSuccess && rc == PYGW_OK) rc = PYGW_E_HIP;
  }
  return rc;
}

extern "C" int pygw_eps_inverse_workspace(int pb, int n_om, size_t* bytes) {
  if (pb <= 0 || n_om <= 0 || !bytes) return PYGW_E_ARG;
  *bytes = eps_inverse_ws_bytes(pb, n_om);
  return PYGW_OK;
}

extern "C" int pygw_eps_inverse(void* eps, int pb, int n_om, void* workspace, size_t ws_bytes, int* info,
                                void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (!eps || pb <= 0 || n_om <= 0) return PYGW_E_ARG;
  const size_t need = eps_inverse_ws_bytes(pb, n_om);
  void* ws = workspace;
  bool own = false;
  if (!ws) {
    HIP_TRY(hipMallocAsync(&ws, need, stream));
    own = true;
  } else if (ws_bytes < need) {
    return PYGW_E_ARG;
  }
  int* dinfo = (int*)((char*)ws + need - 256 - sizeof(int) * (size_t)n_om);
  hipLaunchKernelGGL(pygw_fill_i32, dim3(ceil_div(n_om, 256)), dim3(256), 0, stream, dinfo, n_om, 0);
  int rc = eps_inverse_launch(eps, pb, n_om, ws, dinfo, stream);
  int* h = nullptr;
  if (rc == PYGW_OK && info && !(h = (int*)malloc(sizeof(int) * (size_t)n_om))) rc = PYGW_E_HIP;
  if (rc == PYGW_OK && info) {
    hipError_t e1 = hipMemcpyAsync(h, dinfo, sizeof(int) * (size_t)n_om, hipMemcpyDeviceToHost, stream);
    hipError_t e2 = hipStreamSynchronize(stream);
    int bad = 0;
    for (int w = 0; w < n_om; w++) bad |= h[w];
    if (e1 != hipSuccess || e2 != hipSuccess) rc = PYGW_E_HIP;
    else {
      info[0] = bad;
      if (bad) rc = PYGW_E_NOTPD;
    }
  }
  free(h);
  if (own) {
    hipError_t e = hipFreeAsync(ws, stream);
    if (e != hipSuccess && rc == PYGW_OK) rc = PYGW_E_HIP;
  }
  return rc;
}

// Times the chi0 tile schedules (1, 2, 4 frequencies per block) on the given
// inputs, records the fastest for (device, precision, pb, n_pairs, n_om) and
// returns it in *best (times in ms in times[3] if given).  eps is scratch.
// Split precision: the autotuners consider fold intervals 64 and 128 chunks
// only.  Folding every 256 chunks (s = 4) doubles the FP32 partial-sum length
// and measured 5.9e-7 .. 7.1e-7 relative error per frequency against the
// 1e-6 bar of the FP64 contraction, for about 1-3 % less time; callers may
// still select it explicitly.
static const int kMixedMaxAuto = 2;

extern "C" int pygw_chi0_autotune(const void* M, long stride_k, int n_pairs, int pb, const double* kcw, int n_om,
                                  int k0, int n_k, double coef, int precision, void* eps, int* best, float* times,
                                  void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (!M || !kcw || !eps || !best || n_pairs <= 0 || pb <= 0 || n_om <= 0 || n_k <= 0 || k0 < 0) return PYGW_E_ARG;
  if (precision != 0 && precision != 1) return PYGW_E_ARG;
  void* ws = nullptr;
  const size_t wsb = precision == 1 ? chi0_mixed_ws_bytes(pb, n_pairs, n_om, n_k) : 0;
  if (precision == 1) HIP_TRY(hipMallocAsync(&ws, wsb, stream));
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  const int cand[3] = {1, 2, 4};
  float tbest = 1e30f;
  int nbest = 2, rc = PYGW_OK;
  for (int i = 0; i < 3 && rc == PYGW_OK; i++) {
    const int nw = cand[i];
    if (precision == 1 && nw > kMixedMaxAuto) {
      if (times) times[i] = -1.f;  // not a candidate (accuracy), still selectable by hand
      continue;
    }
    for (int rep = 0; rep < 2 && rc == PYGW_OK; rep++) {  // first run warms up
      if (hipEventRecord(e0, stream) != hipSuccess) { rc = PYGW_E_HIP; break; }
      if (precision == 0) rc = chi0_f64_launch(M, stride_k, n_pairs, pb, kcw, n_om, k0, n_k, coef, eps, 0, nw, stream);
      else rc = chi0_mixed_launch(M, stride_k, n_pairs, pb, kcw, n_om, k0, n_k, coef, 64 * nw, 0, ws, wsb, eps, stream);
      if (hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) { rc = PYGW_E_HIP; break; }
    }
    float ms = 0.f;
    if (rc == PYGW_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = PYGW_E_HIP;
    if (times) times[i] = ms;
    if (rc == PYGW_OK && ms < tbest) { tbest = ms; nbest = nw; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ws) (void)hipFreeAsync(ws, stream);
  if (rc != PYGW_OK) return rc;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[std::make_tuple(dev, precision, pb, n_pairs, n_om)] = nbest;
  }
  *best = nbest;
  return PYGW_OK;
}

// Sets the cached schedule for one shape (nw = 1, 2 or 4; 0 removes it).
extern "C" int pygw_chi0_set_schedule(int precision, int pb, int n_pairs, int n_om, int nw) {
  if ((precision != 0 && precision != 1) || (nw != 0 && nw != 1 && nw != 2 && nw != 4)) return PYGW_E_ARG;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto key = std::make_tuple(dev, precision, pb, n_pairs, n_om);
  if (nw == 0) g_tune.erase(key);
  else g_tune[key] = nw;
  return PYGW_OK;
}

// --- batched split-precision contraction over many q (one launch)
namespace {
bool batch_args_ok(int n_q, const int* pb, const int* n_pairs, int n_om) {
  if (n_q <= 0 || !pb || !n_pairs || n_om <= 0 || n_om > 65535) return false;
  for (int q = 0; q < n_q; q++) {
    if (pb[q] <= 0 || n_pairs[q] <= 0) return false;
    if ((size_t)(n_pairs[q] + 32) * pb[q] * 8 >= ((size_t)1 << 31)) return false;  // 32-bit buffer offsets
  }
  return true;
}
std::mutex g_batch_mu;
std::map<const void*, int> g_batch_sched;  // plan -> schedule chosen by pygw_chi0_batch_autotune
}  // namespace

extern "C" int pygw_chi0_batch_workspace(int n_q, const int* pb, const int* n_pairs, int n_om, size_t* bytes) {
  if (!bytes || !batch_args_ok(n_q, pb, n_pairs, n_om)) return PYGW_E_ARG;
  *bytes = chi0_batch_ws_bytes(n_q, pb, n_pairs, n_om);
  return PYGW_OK;
}

extern "C" int pygw_chi0_batch_plan(int n_q, const void* const* M, const int* pb, const int* n_pairs,
                                    const double* const* kcw, int n_om, double coef, void* const* eps,
                                    void* workspace, size_t ws_bytes, void* stream, void** plan) {
  if (!M || !kcw || !eps || !workspace || !plan || !batch_args_ok(n_q, pb, n_pairs, n_om)) return PYGW_E_ARG;
  for (int q = 0; q < n_q; q++)
    if (!M[q] || !kcw[q] || !eps[q]) return PYGW_E_ARG;
  Chi0BatchPlan* P = nullptr;
  const int rc = chi0_batch_plan(n_q, M, pb, n_pairs, kcw, n_om, coef, eps, workspace, ws_bytes,
                                 (hipStream_t)stream, &P);
  if (rc == PYGW_OK) *plan = P;
  return rc;
}

extern "C" int pygw_chi0_batch_run(void* plan, int schedule, void* stream) {
  if (!plan || (schedule != 0 && schedule != 1 && schedule != 2 && schedule != 4)) return PYGW_E_ARG;
  if (schedule == 0) {
    std::lock_guard<std::mutex> lk(g_batch_mu);
    auto it = g_batch_sched.find(plan);
    schedule = it == g_batch_sched.end() ? 2 : it->second;
  }
  return chi0_batch_run((const Chi0BatchPlan*)plan, 64 * schedule, (hipStream_t)stream);
}

// Times the batch with each schedule s = 1, 2, 4 (FP32 tiles folded into
// FP64 every 64*s chunks) and keeps the fastest for later runs of this plan.
extern "C" int pygw_chi0_batch_autotune(void* plan, int* best, float* times, void* stream_) {
  if (!plan || !best) return PYGW_E_ARG;
  hipStream_t stream = (hipStream_t)stream_;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  const int cand[3] = {1, 2, 4};
  float tbest = 1e30f;
  int nbest = 2, rc = PYGW_OK;
  for (int i = 0; i < 3 && rc == PYGW_OK; i++) {
    if (cand[i] > kMixedMaxAuto) {
      if (times) times[i] = -1.f;
      continue;
    }
    for (int rep = 0; rep < 2 && rc == PYGW_OK; rep++) {
      if (hipEventRecord(e0, stream) != hipSuccess) { rc = PYGW_E_HIP; break; }
      rc = chi0_batch_run((const Chi0BatchPlan*)plan, 64 * cand[i], stream);
      if (rc == PYGW_OK && (hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
        rc = PYGW_E_HIP;
    }
    float ms = 0.f;
    if (rc == PYGW_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = PYGW_E_HIP;
    if (times) times[i] = ms;
    if (rc == PYGW_OK && ms < tbest) { tbest = ms; nbest = cand[i]; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != PYGW_OK) return rc;
  {
    std::lock_guard<std::mutex> lk(g_batch_mu);
    g_batch_sched[plan] = nbest;
  }
  *best = nbest;
  return PYGW_OK;
}

extern "C" int pygw_chi0_batch_free(void* plan) {
  if (!plan) return PYGW_E_ARG;
  {
    std::lock_guard<std::mutex> lk(g_batch_mu);
    g_batch_sched.erase(plan);
  }
  chi0_batch_free((Chi0BatchPlan*)plan);
  return PYGW_OK;
}
