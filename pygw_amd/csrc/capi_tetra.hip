// C-ABI for the tetrahedron-weight kernel (declared in include/pygw_hip.h).
#include "common.h"
#include "../../include/pygw_hip.h"

__global__ void pygw_tw_incidence(const int* tetc, int ntet, int nk, int* inc, int* cnt, int maxinc);
__global__ void pygw_tw_table(const double* enk, int nbnd, const int* kqid, int ncg, int numin, int no, int ne,
                              const double* omega, int nom, int nk, int q0, int pcn, int pp, int layout, double2* tab,
                              double* dtab);
__global__ void pygw_tw_tet(const double* enk, int nbnd, double ef, const int* kqid, int ncg, int numin, int no, int ne,
                            const double* omega, int nom, const int* tetc, int nk, int ntet, int q0, int pcn, int pp,
                            int compat, int layout, const double* dtab, double* w1t, int2* wlist, int* wcount,
                            int2* glist, int* gcount, int* dstat);
__global__ void pygw_tw_whole(const double* enk, int nbnd, const int* kqid, int ncg, int numin, int no, int ne,
                              const double* omega, int nom, const int* tetc, int nk, int q0, int pp, int compat,
                              int layout, const double2* tab, const int2* wlist, const int* wcount, double* w1t,
                              int* dstat);
__global__ void pygw_tw_general(const double* enk, int nbnd, double ef, const int* kqid, int ncg, int numin, int no,
                                int ne, const double* omega, int nom, const int* tetc, int nk, int q0, int pp,
                                int compat, int layout, const int2* glist, const int* gcount, double* w1t, int* dstat);
__global__ void pygw_tw_sum(const double* w1t, const int* inc, const int* cnt, int maxinc, int nk, int ntet, int nom,
                            int q0, int pcn, int pp, int np, int layout, double* kcw, int* dstat);

__global__ void pygw_tw_project(const double* kcw, int np, int nom, int nk, const double* dUl, int nl, int layout,
                                int no, int ne, double* out);

static const int kMaxInc = 24;
static const size_t kW1tBudget = (size_t)8 << 30;  // preferred corner-weight buffer (C5: 2 chunks of pairs)

// head: incidence lists, counts, status words, list counter
static size_t tw_head_bytes(int n_k) {
  size_t b = (size_t)n_k * kMaxInc * sizeof(int) + (size_t)n_k * sizeof(int) + 4 * sizeof(int);
  return (b + 255) & ~(size_t)255;
}

// per band pair of a chunk: two list slots (int2 per tetrahedron),
// w1t[ntet][4][nom] doubles, the table tab[nk][nom] (double2) and dtab[nk]
static size_t tw_pair_bytes(int n_tet, int n_om) {
  const size_t nk = (size_t)n_tet / 6;
  return (size_t)n_tet * 8 * (2 + 4 * (size_t)n_om) + nk * (16 * (size_t)n_om + 8);
}

// pairs per chunk (buffers at pitch pc + kPad) that fit `avail` bytes: all
// of them, else a multiple of 64
static const int kPad = 8;  // the pitch is not a power of two: rows spread over the HBM channels
static int tw_chunk(size_t avail, int n_tet, int n_om, int np) {
  size_t pc = avail / tw_pair_bytes(n_tet, n_om);
  pc = pc > (size_t)kPad ? pc - kPad : 0;
  const size_t lim = ((size_t)1 << 30) / (size_t)n_tet;  // list slots and lanes stay below 2^30
  if (pc > lim) pc = lim;
  if (pc >= (size_t)np) return np;
  return (int)(pc & ~(size_t)63);
}

static size_t tw_ws_bytes(int n_k, int np, int n_om) {
  const int ntet = 6 * n_k;
  int pc = tw_chunk(kW1tBudget, ntet, n_om, np);
  if (pc < 64) pc = np < 64 ? np : 64;
  return tw_head_bytes(n_k) + (size_t)(pc + kPad) * tw_pair_bytes(ntet, n_om);
}

__global__ void pygw_fill_i32(int* __restrict__ x, int n, int v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = v;
}

extern "C" int pygw_version(void) { return 100; }

extern "C" int pygw_tetra_weights_workspace(int n_k, int n_pairs, int n_om, size_t* bytes) {
  if (n_k <= 0 || n_pairs <= 0 || n_om <= 0 || !bytes) return PYGW_E_ARG;
  *bytes = tw_ws_bytes(n_k, n_pairs, n_om);
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
  if ((long)no * ne > (1L << 30) || (long)n_k > 65535 || n_om > 65535) return PYGW_E_ARG;
  const int np = no * ne;
  void* ws = workspace;
  bool own = false;
  size_t have = ws_bytes;
  if (!ws) {
    have = tw_ws_bytes(n_k, np, n_om);
    HIP_TRY(hipMallocAsync(&ws, have, stream));
    own = true;
  }
  const size_t head = tw_head_bytes(n_k);
  const int pc = have > head ? tw_chunk(have - head, n_tet, n_om, np) : 0;
  if (pc <= 0) {  // below one chunk of 64 pairs (or of all pairs)
    if (own) HIP_TRY(hipFreeAsync(ws, stream));
    return PYGW_E_ARG;
  }
  int* inc = (int*)ws;
  int* cnt = inc + (size_t)n_k * kMaxInc;
  int* dstat = cnt + n_k;
  int* gcount = dstat + 2;
  int* wcount = dstat + 3;
  const int pp = pc + kPad;
  double2* tab = (double2*)((char*)ws + head);
  double* w1t = (double*)(tab + (size_t)n_k * n_om * pp);
  double* dtab = w1t + (size_t)n_tet * 4 * n_om * pp;
  int2* glist = (int2*)(dtab + (size_t)n_k * pp);
  int2* wlist = glist + (size_t)n_tet * pp;
  hipLaunchKernelGGL(pygw_fill_i32, dim3(1), dim3(64), 0, stream, dstat, 2, 0);
  hipLaunchKernelGGL(pygw_tw_incidence, dim3(n_k), dim3(64), 0, stream, tetc, n_tet, n_k, inc, cnt, kMaxInc);
  for (int q0 = 0; q0 < np; q0 += pc) {
    const int pcn = np - q0 < pc ? np - q0 : pc;
    hipLaunchKernelGGL(pygw_fill_i32, dim3(1), dim3(64), 0, stream, gcount, 2, 0);
    hipLaunchKernelGGL(pygw_tw_table, dim3(ceil_div(pcn, 64), n_k), dim3(64), 0, stream, enk, n_bnd, kqid, ncg, numin,
                       no, ne, omega, n_om, n_k, q0, pcn, pp, layout, tab, dtab);
    hipLaunchKernelGGL(pygw_tw_tet, dim3((unsigned)n_tet * (unsigned)ceil_div(pcn, 64)), dim3(64), 0, stream, enk,
                       n_bnd, ef, kqid, ncg, numin, no, ne, omega, n_om, tetc, n_k, n_tet, q0, pcn, pp, compat,
                       layout, dtab, w1t, wlist, wcount, glist, gcount, dstat);
    const long gl = 4L * n_tet * pcn;
    hipLaunchKernelGGL(pygw_tw_whole, dim3(ceil_div(gl, 64) < 4096 ? ceil_div(gl, 64) : 4096), dim3(64), 0, stream,
                       enk, n_bnd, kqid, ncg, numin, no, ne, omega, n_om, tetc, n_k, q0, pp, compat, layout, tab,
                       wlist, wcount, w1t, dstat);
    hipLaunchKernelGGL(pygw_tw_general, dim3(ceil_div(gl, 64) < 2048 ? ceil_div(gl, 64) : 2048), dim3(64), 0,
                       stream, enk, n_bnd, ef, kqid, ncg, numin, no, ne, omega, n_om, tetc, n_k, q0, pp, compat,
                       layout, glist, gcount, w1t, dstat);
    hipLaunchKernelGGL(pygw_tw_sum, dim3(ceil_div(pcn, 64), n_om, n_k), dim3(64), 0, stream, w1t, inc, cnt, kMaxInc,
                       n_k, n_tet, n_om, q0, pcn, pp, np, layout, kcw, dstat);
  }
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
