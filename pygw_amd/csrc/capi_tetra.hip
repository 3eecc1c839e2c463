// C-ABI for the tetrahedron-weight kernel (declared in include/pygw_hip.h).
#include "common.h"
#include "../../include/pygw_hip.h"

__global__ void pygw_tw_incidence(const int* tetc, int ntet, int nk, int* inc, int* cnt, int maxinc);
__global__ void pygw_tw_gather(const double* enk, int nbnd, double ef, const int* kqid, int ncg, int numin, int no,
                               int ne, const double* omega, int nom, const int* tetc, int nk, int ntet, const int* inc,
                               const int* cnt, int maxinc, int compat, int layout, double* kcw, int* dstat);

__global__ void pygw_tw_project(const double* kcw, int np, int nom, int nk, const double* dUl, int nl, int layout,
                                int no, int ne, double* out);

static const int kMaxInc = 24;

static size_t tw_ws_bytes(int n_k) {
  size_t b = (size_t)n_k * kMaxInc * sizeof(int) + (size_t)n_k * sizeof(int) + 4 * sizeof(int);
  return (b + 255) & ~(size_t)255;
}

__global__ void pygw_fill_i32(int* __restrict__ x, int n, int v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = v;
}

extern "C" int pygw_version(void) { return 100; }

extern "C" int pygw_tetra_weights_workspace(int n_k, size_t* bytes) {
  if (n_k <= 0 || !bytes) return PYGW_E_ARG;
  *bytes = tw_ws_bytes(n_k);
  return PYGW_OK;
}

extern "C" int pygw_tetra_weights(const double* enk, double ef, const int* kqid, const int* tetc, int ncg, int nomx,
                                  int numin, const double* omega, int n_om, int n_k, int n_bnd, int n_tet, int layout,
                                  int compat, double* kcw, void* workspace, size_t ws_bytes, int* info,
                                  void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  const int no = ncg + nomx + 1, ne = n_bnd - ncg - numin;
  if (!enk || !kqid || !tetc || !omega || !kcw) return PYGW_E_ARG;
  if (no <= 0 || ne <= 0 || n_om <= 0 || n_k <= 0 || n_tet != 6 * n_k || ncg < 0 || numin < 0) return PYGW_E_ARG;
  if (no > n_bnd || ncg + numin >= n_bnd) return PYGW_E_ARG;
  if (layout != 0 && layout != 1) return PYGW_E_ARG;
  if ((long)no * ne > (1L << 30) || (long)n_k > 65535) return PYGW_E_ARG;
  const size_t need = tw_ws_bytes(n_k);
  void* ws = workspace;
  bool own = false;
  if (!ws) {
    HIP_TRY(hipMallocAsync(&ws, need, stream));
    own = true;
  } else if (ws_bytes < need) {
    return PYGW_E_ARG;
  }
  int* inc = (int*)ws;
  int* cnt = inc + (size_t)n_k * kMaxInc;
  int* dstat = cnt + n_k;
  hipLaunchKernelGGL(pygw_fill_i32, dim3(1), dim3(64), 0, stream, dstat, 2, 0);
  hipLaunchKernelGGL(pygw_tw_incidence, dim3(ceil_div(n_k, 256)), dim3(256), 0, stream, tetc, n_tet, n_k, inc, cnt,
                     kMaxInc);
  const int np = no * ne;
  dim3 grid(ceil_div(np, 256), n_k);
  hipLaunchKernelGGL(pygw_tw_gather, grid, dim3(256), 0, stream, enk, n_bnd, ef, kqid, ncg, numin, no, ne, omega,
                     n_om, tetc, n_k, n_tet, inc, cnt, kMaxInc, compat, layout, kcw, dstat);
  HIP_TRY(hipGetLastError());
  int rc = PYGW_OK;
  if (info) {
    int h[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(h, dstat, 2 * sizeof(int), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    info[0] = h[0];
    info[1] = h[1];
    if (h[0] & 1) rc = PYGW_E_GEOM;
  }
  if (own) HIP_TRY(hipFreeAsync(ws, stream));
  return rc;
}

extern "C" int pygw_tetra_weights_svd(const double* kcw_native, int n_k, int n_om, int n_o, int n_e, const double* dUl,
                                      int n_l, int layout, double* out, void* stream) {
  if (!kcw_native || !dUl || !out || n_k <= 0 || n_om <= 0 || n_o <= 0 || n_e <= 0 || n_l <= 0) return PYGW_E_ARG;
  if ((layout != 0 && layout != 1) || n_k > 65535) return PYGW_E_ARG;
  const int np = n_o * n_e;
  hipLaunchKernelGGL(pygw_tw_project, dim3(ceil_div(np, 256), n_k), dim3(256), 0, (hipStream_t)stream, kcw_native,
                     np, n_om, n_k, dUl, n_l, layout, n_o, n_e, out);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}
