// q = 0 (Gamma point) head and wings of PyGW's polarization path, FP64:
//
//   wings (calc_eps, pygw.py:3177-3193):
//     epsw1[I, w] += sum_k sum_p coef*kcw[k][w][p] * M[k][p][I] * pm[k][p]
//   inversion (pygw.py:3202-3263), after (eps + eps^H)/2 + 1 is inverted to A:
//     bw1 = A epsw1, w2b = conj(A epsw2) with epsw2 = conj(epsw1),
//     ww = sum_I epsw2[I] bw1[I], emac = (1 + head - ww, 1 + head),
//     h = 1/(1 + head - ww), epsw1 = -h bw1, epsw2 = -h w2b,
//     A += h bw1 w2b^T, head = h - 1, then A - 1
//   MWM (calc_selfc, pygw.py:3359-3364), rows (b, b + ncg_c) of every k:
//     mwm[k][w][b][b + ncg_c] += coefs2 head[w] + coefs1 sum_I M[I] epsw2[I, w]
//                                               + coefs1 sum_I conj(M[I]) epsw1[I, w]
// epsw1, epsw2 are Fortran [PB, n_om] (PyGW's epsw1[:, iom]).
#include "common.h"
#include "../../include/pygw_hip.h"

int eps_inverse_gj(void* eps, int n, int nom, void* ws, int* dinfo, hipStream_t stream);
int eps_minus_identity(void* eps, int n, int nom, hipStream_t stream);
size_t eps_inverse_ws_bytes(int n, int nom);

typedef double d4 __attribute__((ext_vector_type(4)));

namespace gm {
constexpr int KC = 16;   // pairs staged per step
}

__device__ __forceinline__ double2 ld2(const double2* p) { return *p; }
__device__ __forceinline__ double2 ld2(const float2* p) {
  const float2 v = *p;
  return double2{(double)v.x, (double)v.y};
}

// Wings as a GEMM on FP64 MFMA: the 64 x 64 (I, w) tile of sum_p
// M[p][I] X[p][w], X = coef * kcw * pm, over the k-points of slab z
// (grid.z); 4 waves of 32 x 32 (2 x 2 v_mfma_f64_16x16x4f64 tiles).  The
// partial sums go to part[z] and are added in slab order by
// pygw_gm_wings_sum, so the result does not depend on scheduling.
template <typename CT>
__global__ void __launch_bounds__(256) pygw_gm_wings(const CT* __restrict__ M, long strideK, int np, int pb,
                                                     const double* __restrict__ kcw, int nom, int k0, int nk, int kps,
                                                     double coef, const double2* __restrict__ pm, double2* __restrict__ part) {
  using namespace gm;
  __shared__ double Ar[KC][64 + 1], Ai[KC][64 + 1];  // [p][I]
  __shared__ double Br[KC][64 + 1], Bi[KC][64 + 1];  // [p][w]
  const int I0 = blockIdx.x * 64, W0 = blockIdx.y * 64, z = blockIdx.z;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wi = wave & 1, wj = wave >> 1, li = lane & 15, lk = lane >> 4;
  d4 sr[2][2], si[2][2];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) { sr[x][y] = d4{0, 0, 0, 0}; si[x][y] = d4{0, 0, 0, 0}; }
  const int ka = z * kps, kb = ka + kps < nk ? ka + kps : nk;
  for (int k = ka; k < kb; k++) {
    const CT* Mk = M + (size_t)k * strideK;
    const double* wk = kcw + (size_t)(k0 + k) * nom * np;
    const double2* pk = pm + (size_t)k * np;
    for (int p0 = 0; p0 < np; p0 += KC) {
      __syncthreads();  // the previous step's operands are consumed
      for (int e = tid; e < 64 * KC; e += 256) {
        const int i = e & 63, c = e >> 6, p = p0 + c, I = I0 + i;
        const double2 v = (p < np && I < pb) ? ld2(Mk + (size_t)p * pb + I) : double2{0.0, 0.0};
        Ar[c][i] = v.x;
        Ai[c][i] = v.y;
      }
      for (int e = tid; e < KC * 64; e += 256) {
        const int c = e % KC, j = e / KC, p = p0 + c, w = W0 + j;
        double xr = 0.0, xi = 0.0;
        if (p < np && w < nom) {
          const double t = coef * wk[(size_t)w * np + p];
          const double2 q = pk[p];
          xr = t * q.x;
          xi = t * q.y;
        }
        Br[c][j] = xr;
        Bi[c][j] = xi;
      }
      __syncthreads();
#pragma unroll
      for (int ks = 0; ks < KC / 4; ks++) {
        const int kk = 4 * ks + lk;
        double ar[2], ai[2], nai[2], br[2], bi[2];
#pragma unroll
        for (int t = 0; t < 2; t++) {
          ar[t] = Ar[kk][32 * wi + 16 * t + li];
          ai[t] = Ai[kk][32 * wi + 16 * t + li];
          nai[t] = -ai[t];
          br[t] = Br[kk][32 * wj + 16 * t + li];
          bi[t] = Bi[kk][32 * wj + 16 * t + li];
        }
#pragma unroll
        for (int x = 0; x < 2; x++)
#pragma unroll
          for (int y = 0; y < 2; y++) {
            sr[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[x], br[y], sr[x][y], 0, 0, 0);
            sr[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[x], bi[y], sr[x][y], 0, 0, 0);
            si[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[x], bi[y], si[x][y], 0, 0, 0);
            si[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[x], br[y], si[x][y], 0, 0, 0);
          }
      }
    }
  }
  // D row = 4v + lane/16 -> I, D column = lane%16 -> w
  double2* out = part + (size_t)z * pb * nom;
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int I = I0 + 32 * wi + 16 * x + 4 * v + lk, w = W0 + 32 * wj + 16 * y + li;
        if (I < pb && w < nom) out[I + (size_t)pb * w] = double2{sr[x][y][v], si[x][y][v]};
      }
}

__global__ void pygw_gm_wings_sum(const double2* __restrict__ part, int nz, long n, int beta,
                                  double2* __restrict__ epsw1) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= n) return;
  double2 s = beta ? epsw1[e] : double2{0.0, 0.0};
  for (int z = 0; z < nz; z++) {
    const double2 v = part[(size_t)z * n + e];
    s.x += v.x;
    s.y += v.y;
  }
  epsw1[e] = s;
}

// bw1 = A epsw1 and y2 = A conj(epsw1) per frequency; 64 rows per block,
// four column groups reduced through LDS.
__global__ void __launch_bounds__(256) pygw_gm_mv(const double2* __restrict__ A, int n,
                                                  const double2* __restrict__ w1, double2* __restrict__ bw1,
                                                  double2* __restrict__ y2) {
  __shared__ double2 r1[4][64], r2[4][64];
  const int w = blockIdx.y, g = threadIdx.x >> 6, il = threadIdx.x & 63;
  const int i = blockIdx.x * 64 + il;
  const double2* a = A + (size_t)w * n * n;
  const double2* x = w1 + (size_t)w * n;
  double2 s1 = {0.0, 0.0}, s2 = {0.0, 0.0};
  if (i < n)
    for (int j = g; j < n; j += 4) {
      const double2 aij = a[i + (size_t)n * j], xj = x[j];
      s1.x += aij.x * xj.x - aij.y * xj.y;
      s1.y += aij.x * xj.y + aij.y * xj.x;
      s2.x += aij.x * xj.x + aij.y * xj.y;  // times conj(x_j)
      s2.y += aij.y * xj.x - aij.x * xj.y;
    }
  r1[g][il] = s1;
  r2[g][il] = s2;
  __syncthreads();
  if (g == 0 && i < n) {
    for (int q = 1; q < 4; q++) {
      s1.x += r1[q][il].x; s1.y += r1[q][il].y;
      s2.x += r2[q][il].x; s2.y += r2[q][il].y;
    }
    bw1[i + (size_t)n * w] = s1;
    y2[i + (size_t)n * w] = s2;
  }
}

__device__ __forceinline__ double2 gm_cdiv(double2 a, double2 b) {
  if (fabs(b.x) >= fabs(b.y)) {
    const double r = b.y / b.x, d = b.x + b.y * r;
    return double2{(a.x + a.y * r) / d, (a.y - a.x * r) / d};
  }
  const double r = b.x / b.y, d = b.x * r + b.y;
  return double2{(a.x * r + a.y) / d, (a.y * r - a.x) / d};
}

// one block per frequency: ww, emac, the new head, both wings, w2b
__global__ void __launch_bounds__(256) pygw_gm_head(int n, double2* __restrict__ head, double2* __restrict__ w1,
                                                    double2* __restrict__ w2, const double2* __restrict__ bw1,
                                                    const double2* __restrict__ y2, double2* __restrict__ emac,
                                                    double2* __restrict__ hbuf, double2* __restrict__ w2b) {
  __shared__ double2 red[256];
  __shared__ double2 hs;
  const int w = blockIdx.x, t = threadIdx.x;
  const size_t o = (size_t)n * w;
  double2 s = {0.0, 0.0};
  for (int i = t; i < n; i += 256) {  // epsw2 = conj(epsw1): sum conj(w1) bw1
    const double2 a = w1[o + i], b = bw1[o + i];
    s.x += a.x * b.x + a.y * b.y;
    s.y += a.x * b.y - a.y * b.x;
  }
  red[t] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st) { red[t].x += red[t + st].x; red[t].y += red[t + st].y; }
    __syncthreads();
  }
  if (t == 0) {
    const double2 hd = head[w], ww = red[0];
    const double2 den = double2{1.0 + hd.x - ww.x, hd.y - ww.y};
    emac[2 * w] = den;
    emac[2 * w + 1] = double2{1.0 + hd.x, hd.y};
    const double2 h = gm_cdiv(double2{1.0, 0.0}, den);
    hs = h;
    hbuf[w] = h;
    head[w] = double2{h.x - 1.0, h.y};
  }
  __syncthreads();
  const double2 h = hs;
  for (int i = t; i < n; i += 256) {
    const double2 b = bw1[o + i], c = double2{y2[o + i].x, -y2[o + i].y};  // w2b = conj(A epsw2)
    w1[o + i] = double2{-(h.x * b.x - h.y * b.y), -(h.x * b.y + h.y * b.x)};
    w2[o + i] = double2{-(h.x * c.x - h.y * c.y), -(h.x * c.y + h.y * c.x)};
    w2b[o + i] = c;
  }
}

// A += h (bw1 w2b^T), one thread per element (rows fastest)
__global__ void pygw_gm_rank1(double2* __restrict__ A, int n, const double2* __restrict__ hbuf,
                              const double2* __restrict__ bw1, const double2* __restrict__ w2b) {
  const int w = blockIdx.z;
  const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= n || j >= n) return;
  const size_t o = (size_t)n * w;
  const double2 b = bw1[o + i], c = w2b[o + j], h = hbuf[w];
  const double2 t = double2{b.x * c.x - b.y * c.y, b.x * c.y + b.y * c.x};
  double2* a = A + (size_t)w * n * n + i + (size_t)n * j;
  const double2 v = *a;
  *a = double2{v.x + (h.x * t.x - h.y * t.y), v.y + (h.x * t.y + h.y * t.x)};
}

template <typename CT>
__global__ void __launch_bounds__(256) pygw_gm_mwm_head(const CT* __restrict__ Ms, int nb1, int nb2, int pb,
                                                        int ncg_c, const double2* __restrict__ head,
                                                        const double2* __restrict__ w1,
                                                        const double2* __restrict__ w2, int nom, double c1, double c2,
                                                        double2* __restrict__ mwm) {
  const int k = blockIdx.x, w = blockIdx.y, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const double2* e1 = w1 + (size_t)pb * w;
  const double2* e2 = w2 + (size_t)pb * w;
  for (int b = wave; b < nb1; b += 4) {
    const CT* m = Ms + ((size_t)k * nb1 * nb2 + (size_t)b * nb2 + b + ncg_c) * pb;
    double2 s2 = {0.0, 0.0}, s1 = {0.0, 0.0};
    for (int I = lane; I < pb; I += 64) {
      const double2 x = ld2(m + I), a = e2[I], c = e1[I];
      s2.x += x.x * a.x - x.y * a.y;  // M epsw2
      s2.y += x.x * a.y + x.y * a.x;
      s1.x += x.x * c.x + x.y * c.y;  // conj(M) epsw1
      s1.y += x.x * c.y - x.y * c.x;
    }
    for (int off = 32; off > 0; off >>= 1) {
      s2.x += __shfl_xor(s2.x, off, 64); s2.y += __shfl_xor(s2.y, off, 64);
      s1.x += __shfl_xor(s1.x, off, 64); s1.y += __shfl_xor(s1.y, off, 64);
    }
    if (lane == 0) {
      const double2 hd = head[w];
      double2* o = mwm + (((size_t)k * nom + w) * nb1 + b) * nb2 + b + ncg_c;
      const double2 d = double2{c2 * hd.x + c1 * s2.x + c1 * s1.x, c2 * hd.y + c1 * s2.y + c1 * s1.y};
      *o = double2{o->x + d.x, o->y + d.y};
    }
  }
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
static int wings_slabs(int pb, int nom, int nk) {
  const int tiles = ceil_div(pb, 64) * ceil_div(nom, 64);
  int z = ceil_div(1024, tiles);
  if (z > nk) z = nk;
  return z < 1 ? 1 : z;
}

extern "C" int pygw_chi0_wings_workspace(int pb, int n_om, int n_k, size_t* bytes) {
  if (pb <= 0 || n_om <= 0 || n_k <= 0 || !bytes) return PYGW_E_ARG;
  *bytes = sizeof(double2) * (size_t)wings_slabs(pb, n_om, n_k) * pb * n_om;
  return PYGW_OK;
}

extern "C" int pygw_chi0_wings(const void* M, long stride_k, int n_pairs, int pb, int complex64, const double* kcw,
                               int n_om, int k0, int n_k, double coef, const void* pm, void* epsw1, int beta,
                               void* workspace, size_t ws_bytes, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (!M || !kcw || !pm || !epsw1 || n_pairs <= 0 || pb <= 0 || n_om <= 0 || n_k <= 0 || k0 < 0) return PYGW_E_ARG;
  if (stride_k < (long)n_pairs * pb || (complex64 != 0 && complex64 != 1) || n_om > 65535) return PYGW_E_ARG;
  const int nz = wings_slabs(pb, n_om, n_k);
  const size_t need = sizeof(double2) * (size_t)nz * pb * n_om;
  void* ws = workspace;
  bool own = false;
  if (!ws) {
    HIP_TRY(hipMallocAsync(&ws, need, stream));
    own = true;
  } else if (ws_bytes < need) {
    return PYGW_E_ARG;
  }
  const int kps = ceil_div(n_k, nz);
  const dim3 grid(ceil_div(pb, 64), ceil_div(n_om, 64), ceil_div(n_k, kps));
  if (complex64)
    hipLaunchKernelGGL(pygw_gm_wings<float2>, grid, dim3(256), 0, stream, (const float2*)M, stride_k, n_pairs, pb, kcw,
                       n_om, k0, n_k, kps, coef, (const double2*)pm, (double2*)ws);
  else
    hipLaunchKernelGGL(pygw_gm_wings<double2>, grid, dim3(256), 0, stream, (const double2*)M, stride_k, n_pairs, pb,
                       kcw, n_om, k0, n_k, kps, coef, (const double2*)pm, (double2*)ws);
  const long n = (long)pb * n_om;
  hipLaunchKernelGGL(pygw_gm_wings_sum, dim3(ceil_div(n, 256)), dim3(256), 0, stream, (const double2*)ws, (int)grid.z,
                     n, beta, (double2*)epsw1);
  HIP_TRY(hipGetLastError());
  if (own) HIP_TRY(hipFreeAsync(ws, stream));
  return PYGW_OK;
}

static size_t head_extra_bytes(int pb, int nom) { return sizeof(double2) * (3 * (size_t)pb * nom + nom); }

extern "C" int pygw_eps_inverse_head_workspace(int pb, int n_om, size_t* bytes) {
  if (pb <= 0 || n_om <= 0 || !bytes) return PYGW_E_ARG;
  *bytes = eps_inverse_ws_bytes(pb, n_om) + head_extra_bytes(pb, n_om);
  return PYGW_OK;
}

extern "C" int pygw_eps_inverse_head(void* eps, int pb, int n_om, void* head, void* epsw1, void* epsw2, void* emac,
                                     void* workspace, size_t ws_bytes, int* info, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (!eps || !head || !epsw1 || !epsw2 || !emac || pb <= 0 || n_om <= 0 || n_om > 65535) return PYGW_E_ARG;
  const size_t gj = eps_inverse_ws_bytes(pb, n_om), need = gj + head_extra_bytes(pb, n_om);
  void* ws = workspace;
  bool own = false;
  if (!ws) {
    HIP_TRY(hipMallocAsync(&ws, need, stream));
    own = true;
  } else if (ws_bytes < need) {
    return PYGW_E_ARG;
  }
  int* dinfo = (int*)((char*)ws + gj - 256 - sizeof(int) * (size_t)n_om);
  double2* bw1 = (double2*)((char*)ws + gj);
  double2* y2 = bw1 + (size_t)pb * n_om;
  double2* w2b = y2 + (size_t)pb * n_om;
  double2* hbuf = w2b + (size_t)pb * n_om;
  hipLaunchKernelGGL(pygw_fill_i32, dim3(ceil_div(n_om, 256)), dim3(256), 0, stream, dinfo, n_om, 0);
  int rc = eps_inverse_gj(eps, pb, n_om, ws, dinfo, stream);
  if (rc == PYGW_OK) {
    hipLaunchKernelGGL(pygw_gm_mv, dim3(ceil_div(pb, 64), n_om), dim3(256), 0, stream, (const double2*)eps, pb,
                       (const double2*)epsw1, bw1, y2);
    hipLaunchKernelGGL(pygw_gm_head, dim3(n_om), dim3(256), 0, stream, pb, (double2*)head, (double2*)epsw1,
                       (double2*)epsw2, (const double2*)bw1, (const double2*)y2, (double2*)emac, hbuf, w2b);
    hipLaunchKernelGGL(pygw_gm_rank1, dim3(ceil_div(pb, 64), ceil_div(pb, 4), n_om), dim3(256), 0, stream,
                       (double2*)eps, pb, (const double2*)hbuf, (const double2*)bw1, (const double2*)w2b);
    rc = eps_minus_identity(eps, pb, n_om, stream);
  }
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

extern "C" int pygw_mwm_head(const void* M_selfe, int n_k, int nb1, int nb2, int pb, int ncg_c, int complex64,
                             const void* head, const void* epsw1, const void* epsw2, int n_om, double coefs1,
                             double coefs2, void* mwm, void* stream) {
  if (!M_selfe || !head || !epsw1 || !epsw2 || !mwm || n_k <= 0 || nb1 <= 0 || nb2 <= 0 || pb <= 0 || n_om <= 0)
    return PYGW_E_ARG;
  if (ncg_c < 0 || nb1 + ncg_c > nb2 || (complex64 != 0 && complex64 != 1) || n_om > 65535) return PYGW_E_ARG;
  const dim3 grid(n_k, n_om);
  if (complex64)
    hipLaunchKernelGGL(pygw_gm_mwm_head<float2>, grid, dim3(256), 0, (hipStream_t)stream, (const float2*)M_selfe, nb1,
                       nb2, pb, ncg_c, (const double2*)head, (const double2*)epsw1, (const double2*)epsw2, n_om,
                       coefs1, coefs2, (double2*)mwm);
  else
    hipLaunchKernelGGL(pygw_gm_mwm_head<double2>, grid, dim3(256), 0, (hipStream_t)stream, (const double2*)M_selfe,
                       nb1, nb2, pb, ncg_c, (const double2*)head, (const double2*)epsw1, (const double2*)epsw2, n_om,
                       coefs1, coefs2, (double2*)mwm);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}
