// chi0 contraction of calc_eps (pygw.py:3161-3200) on gfx950 MFMA.
//
//   eps[I,J,w] = sum_k sum_p coef*kcw[k,w,p] * M[k,p,I] * conj(M[k,p,J])
//
// a complex weighted Hermitian rank-K update per Matsubara frequency, with
// K = N_k * N_p.  Only the tiles I_t <= J_t are computed; each block writes
// its tile and the conjugate-transposed mirror.  Output layout is PyGW's
// eps array: Fortran [PB, PB, n_om] (element (I,J,w) at I + PB*J + PB^2*w).
//
// Real formulation per (I,J):  with a = M[p,I], b = w_p * M[p,J]
//   Re += a_r b_r + a_i b_i,   Im += a_i b_r - a_r b_i
// i.e. four real products per frequency on v_mfma_f64_16x16x4f64.
#include "common.h"

typedef double d4 __attribute__((ext_vector_type(4)));

namespace c0 {
constexpr int TB = 64;   // tile edge (I and J)
constexpr int KC = 16;   // pairs staged per step

// Decode the upper-triangle tile pair (It <= Jt) of linear index t.
__device__ __forceinline__ void tile_pair(int t, int T, int& It, int& Jt) {
  int i = 0;
  while (t >= T - i) { t -= T - i; i++; }
  It = i;
  Jt = i + t;
}
}  // namespace c0

// MFMA 16x16x4 f64 operand layouts (lane l):
//   A[i][k]: i = l%16, k = l/16;  B[k][j]: j = l%16, k = l/16;
//   D[i][j] (4 values v): j = l%16, i = 4*v + l/16  (not the FP16 order).
template <int NW>
__global__ void __launch_bounds__(256) pygw_chi0_f64(const double2* __restrict__ M, long strideK, int np, int pb,
                                                     const double* __restrict__ kcw, int nom, int nk, double coef,
                                                     int T, double2* __restrict__ eps, int beta) {
  using namespace c0;
  __shared__ double2 As[KC][TB];
  __shared__ double2 Bs[KC][TB];
  __shared__ double Ws[NW][KC];
  int It, Jt;
  tile_pair(blockIdx.x, T, It, Jt);
  const int w0 = blockIdx.y * NW;
  const int I0 = It * TB, J0 = Jt * TB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int li = lane & 15, lk = lane >> 4;

  d4 accR[NW][2][2], accI[NW][2][2];
#pragma unroll
  for (int w = 0; w < NW; w++)
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) {
        accR[w][a][b] = d4{0.0, 0.0, 0.0, 0.0};
        accI[w][a][b] = d4{0.0, 0.0, 0.0, 0.0};
      }

  // chunks c = (k, p0) walked in order; the next chunk's global loads are
  // issued before the MFMA work on the current one (register prefetch)
  const int npc = (np + KC - 1) / KC;
  const long nch = (long)nk * npc;
  constexpr int NQ = (KC * TB) / 256;
  double2 pa[NQ], pv[NQ];
  double pw = 0.0;
  auto fetch = [&](long ch) {
    const int k = (int)(ch / npc), p0 = (int)(ch % npc) * KC;
    const double2* Mk = M + (size_t)k * strideK;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const int e = tid + 256 * q;
      const int r = e / TB, c = e % TB;
      const int p = p0 + r;
      pa[q] = double2{0.0, 0.0};
      pv[q] = double2{0.0, 0.0};
      if (p < np) {
        if (I0 + c < pb) pa[q] = Mk[(size_t)p * pb + I0 + c];
        if (J0 + c < pb) pv[q] = Mk[(size_t)p * pb + J0 + c];
      }
    }
    pw = 0.0;
    if (tid < NW * KC) {
      const int w = tid / KC, r = tid % KC;
      if (p0 + r < np && w0 + w < nom) pw = coef * kcw[((size_t)k * nom + (w0 + w)) * np + p0 + r];
    }
  };
  fetch(0);
  for (long ch = 0; ch < nch; ch++) {
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const int e = tid + 256 * q;
      As[e / TB][e % TB] = pa[q];
      Bs[e / TB][e % TB] = pv[q];
    }
    if (tid < NW * KC) Ws[tid / KC][tid % KC] = pw;
    __syncthreads();
    if (ch + 1 < nch) fetch(ch + 1);
#pragma unroll
    for (int ks = 0; ks < KC / 4; ks++) {
      const int kk = ks * 4 + lk;
      double ar[2], ai[2], br[2], bi[2];
#pragma unroll
      for (int t = 0; t < 2; t++) {
        const double2 a = As[kk][32 * wr + 16 * t + li];
        const double2 b = Bs[kk][32 * wc + 16 * t + li];
        ar[t] = a.x; ai[t] = a.y; br[t] = b.x; bi[t] = b.y;
      }
#pragma unroll
      for (int w = 0; w < NW; w++) {
        const double wv = Ws[w][kk];
        double bwr[2], bwi[2];
#pragma unroll
        for (int t = 0; t < 2; t++) { bwr[t] = br[t] * wv; bwi[t] = bi[t] * wv; }
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
          for (int b = 0; b < 2; b++) {
            accR[w][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], bwr[b], accR[w][a][b], 0, 0, 0);
            accR[w][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], bwi[b], accR[w][a][b], 0, 0, 0);
            accI[w][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], bwr[b], accI[w][a][b], 0, 0, 0);
            accI[w][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ar[a], bwi[b], accI[w][a][b], 0, 0, 0);
          }
      }
    }
    __syncthreads();
  }
  const size_t pb2 = (size_t)pb * pb;
#pragma unroll
  for (int w = 0; w < NW; w++) {
    if (w0 + w >= nom) break;
    double2* E = eps + pb2 * (w0 + w);
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int I = I0 + 32 * wr + 16 * a + 4 * v + lk;
          const int J = J0 + 32 * wc + 16 * b + li;
          if (I < pb && J < pb) {
            const double re = accR[w][a][b][v], im = accI[w][a][b][v];
            double2* e1 = E + (size_t)I + (size_t)pb * J;
            if (beta) { double2 o = *e1; *e1 = double2{o.x + re, o.y + im}; }
            else *e1 = double2{re, im};
            if (It != Jt) {
              double2* e2 = E + (size_t)J + (size_t)pb * I;
              if (beta) { double2 o = *e2; *e2 = double2{o.x + re, o.y - im}; }
              else *e2 = double2{re, -im};
            }
          }
        }
  }
}

template <int NW>
static void chi0_f64_go(const void* M, long strideK, int np, int pb, const double* kcw, int nom, int nk, double coef,
                        void* eps, int beta, hipStream_t stream) {
  const int T = ceil_div(pb, c0::TB);
  dim3 grid(T * (T + 1) / 2, ceil_div(nom, NW));
  hipLaunchKernelGGL(pygw_chi0_f64<NW>, grid, dim3(256), 0, stream, (const double2*)M, strideK, np, pb, kcw, nom,
                     nk, coef, T, (double2*)eps, beta);
}

// nw: frequencies per block (1, 2 or 4) -- the tile schedule the autotuner picks
int chi0_f64_launch(const void* M, long strideK, int np, int pb, const double* kcw, int nom, int k0, int nk,
                    double coef, void* eps, int beta, int nw, hipStream_t stream) {
  kcw += (size_t)k0 * nom * np;
  if (nw == 1) chi0_f64_go<1>(M, strideK, np, pb, kcw, nom, nk, coef, eps, beta, stream);
  else if (nw == 4) chi0_f64_go<4>(M, strideK, np, pb, kcw, nom, nk, coef, eps, beta, stream);
  else chi0_f64_go<2>(M, strideK, np, pb, kcw, nom, nk, coef, eps, beta, stream);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}
