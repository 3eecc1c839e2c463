// C-ABI for the contractions and the dielectric inversion (include/pygw_hip.h).
#include "common.h"

int chi0_f64_launch(const void* M, long strideK, int np, int pb, const double* kcw, int nom, int k0, int nk,
                    double coef, void* eps, int beta, hipStream_t stream);
int chi0_mixed_launch(const void* M, long strideK, int np, int pb, const double* kcw, int nom, int k0, int nkc,
                      double coef, int kflush, int beta, int have_scales, void* ws, void* eps, hipStream_t stream);
size_t chi0_mixed_ws_bytes(int pb, int nom);
int eps_inverse_launch(void* eps, int n, int nom, void* ws, int* dinfo, hipStream_t stream);
size_t eps_inverse_ws_bytes(int n, int nom);

extern "C" int pygw_chi0_workspace(int pb, int n_om, int precision, size_t* bytes) {
  if (pb <= 0 || n_om <= 0 || !bytes || (precision != 0 && precision != 1)) return PYGW_E_ARG;
  *bytes = precision == 0 ? 0 : chi0_mixed_ws_bytes(pb, n_om);
  return PYGW_OK;
}

extern "C" int pygw_chi0_accumulate(const void* M, long stride_k, int n_pairs, int pb, const double* kcw, int n_om,
                                    int k0, int n_k, double coef, int precision, int beta, void* eps,
                                    void* workspace, size_t ws_bytes, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (!M || !kcw || !eps || n_pairs <= 0 || pb <= 0 || n_om <= 0 || n_k <= 0 || k0 < 0) return PYGW_E_ARG;
  if (stride_k < (long)n_pairs * pb || (beta != 0 && beta != 1)) return PYGW_E_ARG;
  if (precision == 0) return chi0_f64_launch(M, stride_k, n_pairs, pb, kcw, n_om, k0, n_k, coef, eps, beta, stream);
  if (precision != 1) return PYGW_E_ARG;
  const size_t need = chi0_mixed_ws_bytes(pb, n_om);
  void* ws = workspace;
  bool own = false;
  if (!ws) {
    HIP_TRY(hipMallocAsync(&ws, need, stream));
    own = true;
  } else if (ws_bytes < need) {
    return PYGW_E_ARG;
  }
  const int steps_per_k = (n_pairs + 31) / 32;
  int kflush = 256 / steps_per_k;
  if (kflush < 1) kflush = 1;
  int rc = chi0_mixed_launch(M, stride_k, n_pairs, pb, kcw, n_om, k0, n_k, coef, kflush, beta, 0, ws, eps, stream);
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
  if (rc == PYGW_OK && info) {
    int* h = (int*)malloc(sizeof(int) * (size_t)n_om);
    if (!h) return PYGW_E_ARG;
    hipError_t e1 = hipMemcpyAsync(h, dinfo, sizeof(int) * (size_t)n_om, hipMemcpyDeviceToHost, stream);
    hipError_t e2 = hipStreamSynchronize(stream);
    int bad = 0;
    for (int w = 0; w < n_om; w++) bad |= h[w];
    free(h);
    if (e1 != hipSuccess || e2 != hipSuccess) rc = PYGW_E_HIP;
    else {
      info[0] = bad;
      if (bad) rc = PYGW_E_NOTPD;
    }
  }
  if (own) {
    hipError_t e = hipFreeAsync(ws, stream);
    if (e != hipSuccess && rc == PYGW_OK) rc = PYGW_E_HIP;
  }
  return rc;
}
