// Split-precision chi0 contraction on gfx950 MFMA (v_mfma_f32_16x16x32_f16).
//
//   eps[I,J,w] += sum_k sum_p coef*kcw[k,w,p] * M[k,p,I] * conj(M[k,p,J])
//
// M (complex64) is scaled per column by an exact power of two s_I (so every
// |Re|,|Im| < 2^14, high in FP16's normal range so that the low halves of
// small entries stay normal too) and the weights per frequency by a power of
// two t_w (|w t_w| < 1); each
// real operand x is split into FP16 x_hi = fp16(x), x_lo = fp16(x - x_hi)
// (22 significant bits), and each real product uses three MFMA terms
//   a_hi b_hi + a_hi b_lo + a_lo b_hi        (FP32 accumulation).
// The FP32 partial sums are folded into the FP64 eps every `kflush` k-points
// (multiplied back by the exact factor 1/(s_I s_J t_w)), so no FP32
// accumulator sums more than kflush*N_p/32 MFMA steps: the compensated
// FP64 sum across k keeps the long K = N_k*N_p reduction in FP64.
//
// The weight enters on the J operand, re-split per frequency:
//   b = (w_p t_w) (s_J M[p,J]),  Re += a_r b_r + a_i b_i,  Im += a_i b_r - a_r b_i.
// Only tiles I_t <= J_t are accumulated (into eps, which must hold the
// running sum, zero for a fresh contraction); pygw_chi0_mirror fills the
// rest from the Hermitian symmetry.
#include "common.h"

typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));

namespace cm {
constexpr int TB = 64;   // block tile edge
constexpr int KC = 32;   // pairs per MFMA step

__device__ __forceinline__ void tile_pair(int t, int T, int& It, int& Jt) {
  int i = 0;
  while (t >= T - i) { t -= T - i; i++; }
  It = i;
  Jt = i + t;
}

__device__ __forceinline__ void split(float x, _Float16& hi, _Float16& lo) {
  hi = (_Float16)x;
  lo = (_Float16)(x - (float)hi);
}
}  // namespace cm

// MFMA 16x16x32 f16 operand layouts (lane l, element j = 0..7 of the h8):
//   A[i][k]: i = l%16, k = 8*(l/16) + j;  B[k][n]: n = l%16, k = 8*(l/16) + j;
//   D[i][n] (4 values v): n = l%16, i = 4*(l/16) + v.
template <int NW>
__global__ void __launch_bounds__(256) pygw_chi0_mixed(const float2* __restrict__ M, long strideK, int np, int pb,
                                                       const double* __restrict__ kcw, int nom, int k0, int nk,
                                                       double coef, const float* __restrict__ cs,
                                                       const double* __restrict__ ws, int kflush, int T,
                                                       double2* __restrict__ eps) {
  using namespace cm;
  // staged tiles: [pair][column] (real, imag), pre-scaled by the column scale
  __shared__ float Ar[KC][TB + 1], Ai[KC][TB + 1], Br[KC][TB + 1], Bi[KC][TB + 1];
  __shared__ float Wt[NW][KC];
  int It, Jt;
  tile_pair(blockIdx.x, T, It, Jt);
  const int w0 = blockIdx.y * NW;
  const int I0 = It * TB, J0 = Jt * TB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int li = lane & 15, lk = lane >> 4;

  f4 aR[NW][2][2], aI[NW][2][2];
  auto zero = [&]() {
#pragma unroll
    for (int w = 0; w < NW; w++)
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
          aR[w][a][b] = f4{0.f, 0.f, 0.f, 0.f};
          aI[w][a][b] = f4{0.f, 0.f, 0.f, 0.f};
        }
  };
  auto flush = [&]() {
    const size_t pb2 = (size_t)pb * pb;
#pragma unroll
    for (int w = 0; w < NW; w++) {
      if (w0 + w >= nom) break;
      const double tw = ws[w0 + w];
      double2* E = eps + pb2 * (w0 + w);
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int I = I0 + 32 * wr + 16 * a + 4 * lk + v;
            const int J = J0 + 32 * wc + 16 * b + li;
            if (I < pb && J < pb) {
              const double f = 1.0 / ((double)cs[I] * (double)cs[J] * tw);  // exact: powers of two
              double2* e = E + (size_t)I + (size_t)pb * J;
              double2 o = *e;
              *e = double2{o.x + (double)aR[w][a][b][v] * f, o.y + (double)aI[w][a][b][v] * f};
            }
          }
    }
  };
  zero();
  // chunks (k, p0) in order, next chunk prefetched into registers
  const int npc = (np + KC - 1) / KC;
  const long nch = (long)nk * npc;
  constexpr int NQ = (KC * TB) / 256;
  float2 pa[NQ], pv[NQ];
  float pw = 0.f;
  auto fetch = [&](long ch) {
    const int kk = (int)(ch / npc), p0 = (int)(ch % npc) * KC;
    const float2* Mk = M + (size_t)kk * strideK;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const int e = tid + 256 * q;
      const int r = e / TB, c = e % TB;
      const int p = p0 + r;
      pa[q] = float2{0.f, 0.f};
      pv[q] = float2{0.f, 0.f};
      if (p < np) {
        if (I0 + c < pb) { pa[q] = Mk[(size_t)p * pb + I0 + c]; pa[q].x *= cs[I0 + c]; pa[q].y *= cs[I0 + c]; }
        if (J0 + c < pb) { pv[q] = Mk[(size_t)p * pb + J0 + c]; pv[q].x *= cs[J0 + c]; pv[q].y *= cs[J0 + c]; }
      }
    }
    pw = 0.f;
    if (tid < NW * KC) {
      const int w = tid / KC, r = tid % KC;
      if (p0 + r < np && w0 + w < nom)
        pw = (float)(coef * kcw[((size_t)(k0 + kk) * nom + (w0 + w)) * np + p0 + r] * ws[w0 + w]);
    }
  };
  fetch(0);
  for (long ch = 0; ch < nch; ch++) {
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const int e = tid + 256 * q;
      const int r = e / TB, c = e % TB;
      Ar[r][c] = pa[q].x; Ai[r][c] = pa[q].y;
      Br[r][c] = pv[q].x; Bi[r][c] = pv[q].y;
    }
    if (tid < NW * KC) Wt[tid / KC][tid % KC] = pw;
    __syncthreads();
    if (ch + 1 < nch) fetch(ch + 1);
    // A operands (I side): hi/lo of a_r, a_i and of -a_r
    h8 arh[2], arl[2], aih[2], ail[2], anh[2], anl[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const int col = 32 * wr + 16 * t + li;
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int r = 8 * lk + j;
        _Float16 h, l;
        cm::split(Ar[r][col], h, l); arh[t][j] = h; arl[t][j] = l; anh[t][j] = -h; anl[t][j] = -l;
        cm::split(Ai[r][col], h, l); aih[t][j] = h; ail[t][j] = l;
      }
    }
    float br[2][8], bi[2][8];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const int col = 32 * wc + 16 * t + li;
#pragma unroll
      for (int j = 0; j < 8; j++) { br[t][j] = Br[8 * lk + j][col]; bi[t][j] = Bi[8 * lk + j][col]; }
    }
#pragma unroll
    for (int w = 0; w < NW; w++) {
      h8 brh[2], brl[2], bih[2], bil[2];
#pragma unroll
      for (int t = 0; t < 2; t++)
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const float wv = Wt[w][8 * lk + j];
          _Float16 h, l;
          cm::split(br[t][j] * wv, h, l); brh[t][j] = h; brl[t][j] = l;
          cm::split(bi[t][j] * wv, h, l); bih[t][j] = h; bil[t][j] = l;
        }
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
          f4 r = aR[w][a][b], m = aI[w][a][b];
          r = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh[a], brh[b], r, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh[a], brl[b], r, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_16x16x32_f16(arl[a], brh[b], r, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih[a], bih[b], r, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih[a], bil[b], r, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_16x16x32_f16(ail[a], bih[b], r, 0, 0, 0);
          m = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih[a], brh[b], m, 0, 0, 0);
          m = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih[a], brl[b], m, 0, 0, 0);
          m = __builtin_amdgcn_mfma_f32_16x16x32_f16(ail[a], brh[b], m, 0, 0, 0);
          m = __builtin_amdgcn_mfma_f32_16x16x32_f16(anh[a], bih[b], m, 0, 0, 0);
          m = __builtin_amdgcn_mfma_f32_16x16x32_f16(anh[a], bil[b], m, 0, 0, 0);
          m = __builtin_amdgcn_mfma_f32_16x16x32_f16(anl[a], bih[b], m, 0, 0, 0);
          aR[w][a][b] = r;
          aI[w][a][b] = m;
        }
    }
    __syncthreads();
    const int kk = (int)(ch / npc);
    if ((ch % npc) == npc - 1 && ((kk + 1) % kflush == 0 || kk + 1 == nk)) {
      flush();
      zero();
    }
  }
}

// Lower-triangle tiles from the upper ones: eps[J,I] = conj(eps[I,J]) for
// I in tile It < Jt.  One thread per element of the strict lower part.
__global__ void pygw_chi0_mirror(double2* __restrict__ eps, int pb, int nom) {
  const long n = (long)pb * pb;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int I = (int)(e % pb), J = (int)(e / pb);  // element (I, J), column-major
  if ((I / cm::TB) <= (J / cm::TB)) return;         // computed tile
  double2* E = eps + (size_t)blockIdx.y * n;
  const double2 v = E[(size_t)J + (size_t)pb * I];
  E[(size_t)I + (size_t)pb * J] = double2{v.x, -v.y};
}

// per-column scale: cs[I] = 2^(14 - ceil(log2(max_{k,p} max(|Re|,|Im|))))
__global__ void pygw_colmax(const float2* __restrict__ M, long strideK, int np, int pb, int nk,
                            unsigned int* __restrict__ cmax) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= pb) return;
  float m = 0.f;
  for (int k = blockIdx.y; k < nk; k += gridDim.y)
    for (int p = 0; p < np; p++) {
      const float2 v = M[(size_t)k * strideK + (size_t)p * pb + I];
      m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
    }
  atomicMax(&cmax[I], __float_as_uint(m));  // non-negative floats order as uints
}

__global__ void pygw_colscale(const unsigned int* __restrict__ cmax, int pb, float* __restrict__ cs) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= pb) return;
  const float m = __uint_as_float(cmax[I]);
  int e = 0;
  if (m > 0.f) { frexpf(m, &e); }  // m = f * 2^e, f in [0.5, 1)
  cs[I] = ldexpf(1.0f, 14 - e);     // |x * cs| < 2^14
}

// per-frequency weight scale: ws[w] = 2^-e, e = exponent of max_{k,p} |coef*kcw|
__global__ void pygw_wmax(const double* __restrict__ kcw, int np, int nom, int nk, double coef,
                          unsigned long long* __restrict__ wmax) {
  const int w = blockIdx.y;
  double m = 0.0;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)nk * np; e += (long)gridDim.x * blockDim.x) {
    const int k = (int)(e / np), p = (int)(e % np);
    m = fmax(m, fabs(coef * kcw[((size_t)k * nom + w) * np + p]));
  }
  atomicMax(&wmax[w], (unsigned long long)__double_as_longlong(m));
}

__global__ void pygw_wscale(const unsigned long long* __restrict__ wmax, int nom, double* __restrict__ ws) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nom) return;
  const double m = __longlong_as_double((long long)wmax[w]);
  int e = 0;
  if (m > 0.0) frexp(m, &e);
  ws[w] = ldexp(1.0, -e);
}

__global__ void pygw_zero_u32(unsigned int* __restrict__ x, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) x[e] = 0u;
}

__global__ void pygw_zero_c128(double2* __restrict__ x, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    x[e] = double2{0.0, 0.0};
}

size_t chi0_mixed_ws_bytes(int pb, int nom) {
  return (size_t)pb * 8 + (size_t)nom * 16 + 512;
}

// M: complex64 [n_kc][np][pb] holding k = k0..k0+n_kc-1; kcw covers all k.
// scales: computed from this M chunk unless `have_scales` (then taken from ws).
int chi0_mixed_launch(const void* M, long strideK, int np, int pb, const double* kcw, int nom, int k0, int nkc,
                      double coef, int kflush, int beta, int have_scales, int nw, void* ws, void* eps,
                      hipStream_t stream) {
  float* cs = (float*)ws;
  unsigned int* cmax = (unsigned int*)(cs + pb);
  double* wsc = (double*)(((uintptr_t)(cmax + pb) + 15) & ~(uintptr_t)15);
  unsigned long long* wmax = (unsigned long long*)(wsc + nom);
  if (!have_scales) {
    hipLaunchKernelGGL(pygw_zero_u32, dim3(1), dim3(256), 0, stream, cmax, (long)pb);
    hipLaunchKernelGGL(pygw_zero_u32, dim3(1), dim3(256), 0, stream, (unsigned int*)wmax, 2L * nom);
    hipLaunchKernelGGL(pygw_colmax, dim3(ceil_div(pb, 64), nkc < 64 ? nkc : 64), dim3(64), 0, stream,
                       (const float2*)M, strideK, np, pb, nkc, cmax);
    hipLaunchKernelGGL(pygw_colscale, dim3(ceil_div(pb, 256)), dim3(256), 0, stream, cmax, pb, cs);
    hipLaunchKernelGGL(pygw_wmax, dim3(64, nom), dim3(256), 0, stream, kcw + (size_t)k0 * nom * np, np, nom, nkc,
                       coef, wmax);
    hipLaunchKernelGGL(pygw_wscale, dim3(ceil_div(nom, 256)), dim3(256), 0, stream, wmax, nom, wsc);
  }
  const long ntot = (long)pb * pb * nom;
  if (!beta) hipLaunchKernelGGL(pygw_zero_c128, dim3(2048), dim3(256), 0, stream, (double2*)eps, ntot);
  const int T = ceil_div(pb, cm::TB);
  if (nw == 1) {
    dim3 grid(T * (T + 1) / 2, nom);
    hipLaunchKernelGGL(pygw_chi0_mixed<1>, grid, dim3(256), 0, stream, (const float2*)M, strideK, np, pb, kcw, nom,
                       k0, nkc, coef, cs, wsc, kflush, T, (double2*)eps);
  } else if (nw == 4) {
    dim3 grid(T * (T + 1) / 2, ceil_div(nom, 4));
    hipLaunchKernelGGL(pygw_chi0_mixed<4>, grid, dim3(256), 0, stream, (const float2*)M, strideK, np, pb, kcw, nom,
                       k0, nkc, coef, cs, wsc, kflush, T, (double2*)eps);
  } else {
    dim3 grid(T * (T + 1) / 2, ceil_div(nom, 2));
    hipLaunchKernelGGL(pygw_chi0_mixed<2>, grid, dim3(256), 0, stream, (const float2*)M, strideK, np, pb, kcw, nom,
                       k0, nkc, coef, cs, wsc, kflush, T, (double2*)eps);
  }
  hipLaunchKernelGGL(pygw_chi0_mirror, dim3(ceil_div((long)pb * pb, 256), nom), dim3(256), 0, stream,
                     (double2*)eps, pb, nom);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}
