// MWM contraction of calc_selfc (pygw.py:3353-3358) on gfx950 MFMA, FP64.
//
//   mwm[w][r] = wkq * sum_J ( sum_I conj(M[r,I]) W[I,J,w] ) * M[r,J]
//
// r runs over all (k, ib1, ib2) rows stacked (n_rows = N_k * nb1 * nb2), W
// is PyGW's eps after the inversion (eps^-1 - 1), Fortran [PB, PB, n_om].
// One block owns 64 rows and one frequency: it sweeps the J tiles, forms the
// 64 x 64 tile of conj(M) W on v_mfma_f64_16x16x4f64 (four real products per
// complex product), multiplies it elementwise by M[r, J] and reduces over J
// in registers and LDS, so mw never leaves the chip and no atomics are used.
#include "common.h"

// No FMA contraction: the FP16 hi/lo splits below must see one rounded f32
// value in both halves.
#pragma clang fp contract(off)

typedef double d4 __attribute__((ext_vector_type(4)));

namespace mw {
constexpr int TR = 64;  // rows per block
constexpr int TJ = 64;  // J tile
constexpr int KC = 16;  // I per stage
}

// out: [N_k][n_om][rows_per_k] (PyGW's mwm[iom, ib1, ib2] per k), row = k*rows_per_k + r
__global__ void __launch_bounds__(256) pygw_mwm_f64(const double2* __restrict__ M, int nrows, int pb,
                                                    const double2* __restrict__ W, int nom, double wkq,
                                                    int rpk, double2* __restrict__ out) {
  using namespace mw;
  __shared__ double2 As[TR][KC + 1];   // M rows (conjugated on use)
  __shared__ double2 Bs[KC][TJ + 1];   // W tile
  __shared__ double2 racc[2][TR];
  const int r0 = blockIdx.x * TR;
  const int w = blockIdx.y;
  const double2* Ww = W + (size_t)w * pb * pb;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int li = lane & 15, lk = lane >> 4;
  for (int e = tid; e < 2 * TR; e += 256) racc[e / TR][e % TR] = double2{0.0, 0.0};
  // per-lane partial row sums: rows 32*wr + 16*a + 4*v + lk
  double2 rs[2][4];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int v = 0; v < 4; v++) rs[a][v] = double2{0.0, 0.0};

  for (int J0 = 0; J0 < pb; J0 += TJ) {
    d4 aR[2][2], aI[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) { aR[a][b] = d4{0, 0, 0, 0}; aI[a][b] = d4{0, 0, 0, 0}; }
    for (int I0 = 0; I0 < pb; I0 += KC) {
      for (int e = tid; e < TR * KC; e += 256) {
        const int r = e / KC, c = e % KC;
        const int row = r0 + r, I = I0 + c;
        As[r][c] = (row < nrows && I < pb) ? M[(size_t)row * pb + I] : double2{0.0, 0.0};
      }
      for (int e = tid; e < KC * TJ; e += 256) {
        const int c = e % KC, j = e / KC;  // consecutive threads: consecutive I (column-major W)
        const int I = I0 + c, J = J0 + j;
        Bs[c][j] = (I < pb && J < pb) ? Ww[(size_t)I + (size_t)pb * J] : double2{0.0, 0.0};
      }
      __syncthreads();
#pragma unroll
      for (int ks = 0; ks < KC / 4; ks++) {
        const int kk = ks * 4 + lk;
        double ar[2], ai[2], br[2], bi[2];
#pragma unroll
        for (int t = 0; t < 2; t++) {
          const double2 x = As[32 * wr + 16 * t + li][kk];
          const double2 y = Bs[kk][32 * wc + 16 * t + li];
          ar[t] = x.x; ai[t] = x.y; br[t] = y.x; bi[t] = y.y;
        }
        // conj(m) * W: Re = mr Wr + mi Wi, Im = mr Wi - mi Wr
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
          for (int b = 0; b < 2; b++) {
            aR[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[b], aR[a][b], 0, 0, 0);
            aR[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], bi[b], aR[a][b], 0, 0, 0);
            aI[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], bi[b], aI[a][b], 0, 0, 0);
            aI[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[a], br[b], aI[a][b], 0, 0, 0);
          }
      }
      __syncthreads();
    }
    // (conj(M) W)[r, J] * M[r, J], summed over this lane's J
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = r0 + 32 * wr + 16 * a + 4 * v + lk;
        double2 s = rs[a][v];
#pragma unroll
        for (int b = 0; b < 2; b++) {
          const int J = J0 + 32 * wc + 16 * b + li;
          if (row < nrows && J < pb) {
            const double2 m = M[(size_t)row * pb + J];
            const double xr = aR[a][b][v], xi = aI[a][b][v];
            s.x += xr * m.x - xi * m.y;
            s.y += xr * m.y + xi * m.x;
          }
        }
        rs[a][v] = s;
      }
  }
  // reduce over the 16 lanes that share rows, then over the two column waves
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      double x = rs[a][v].x, y = rs[a][v].y;
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) {
        x += __shfl_xor(x, off, 64);
        y += __shfl_xor(y, off, 64);
      }
      if (li == 0) racc[wc][32 * wr + 16 * a + 4 * v + lk] = double2{x, y};
    }
  __syncthreads();
  for (int r = tid; r < TR; r += 256) {
    const int row = r0 + r;
    if (row < nrows) {
      const double2 s0 = racc[0][r], s1 = racc[1][r];
      const int k = row / rpk, rr = row % rpk;
      out[((size_t)k * nom + w) * rpk + rr] = double2{wkq * (s0.x + s1.x), wkq * (s0.y + s1.y)};
    }
  }
}

int mwm_f64_launch(const void* M, int nrows, int pb, const void* W, int nom, double wkq, int rpk, void* out,
                   hipStream_t stream) {
  dim3 grid(ceil_div(nrows, mw::TR), nom);
  hipLaunchKernelGGL(pygw_mwm_f64, grid, dim3(256), 0, stream, (const double2*)M, nrows, pb, (const double2*)W, nom,
                     wkq, rpk, (double2*)out);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

// ---------------------------------------------------------------------------
// Split-precision MWM (v_mfma_f32_16x16x32_f16): conj(M) W with M rows
// scaled by alpha_r = 2^(14-e_r) and W columns by beta_J = 2^(14-e_J) (exact
// powers of two, so every operand is < 2^14, high in FP16's normal range and
// the low halves of small entries stay normal), each real operand split into FP16
// hi + lo, three products per real product, FP32 accumulation over I
// (PB/32 MFMA steps), then unscaled, multiplied by M[r,J] and summed over J
// in FP64.
// ---------------------------------------------------------------------------
typedef float f4m __attribute__((ext_vector_type(4)));
typedef _Float16 h8m __attribute__((ext_vector_type(8)));

__device__ __forceinline__ void mw_split(float x, _Float16& hi, _Float16& lo) {
  asm volatile("" : "+v"(x));  // one f32 rounding of x feeds both halves
  hi = (_Float16)x;
  lo = (_Float16)(x - (float)hi);
}

__global__ void __launch_bounds__(256) pygw_mwm_mixed(const float2* __restrict__ M, int nrows, int pb,
                                                      const double2* __restrict__ W, int nom, double wkq, int rpk,
                                                      const float* __restrict__ rs_, const float* __restrict__ cs_,
                                                      double2* __restrict__ out) {
  constexpr int TR = 64, TJ = 64, KC = 32;
  __shared__ float Ar[TR][KC + 1], Ai[TR][KC + 1];
  __shared__ float Br[KC][TJ + 1], Bi[KC][TJ + 1];
  __shared__ double2 racc[2][TR];
  const int r0 = blockIdx.x * TR;
  const int w = blockIdx.y;
  const double2* Ww = W + (size_t)w * pb * pb;
  const float* cs = cs_ + (size_t)w * pb;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int li = lane & 15, lk = lane >> 4;
  double2 rs[2][4];
  for (int a = 0; a < 2; a++)
    for (int v = 0; v < 4; v++) rs[a][v] = double2{0.0, 0.0};
  for (int J0 = 0; J0 < pb; J0 += TJ) {
    f4m aR[2][2], aI[2][2];
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) { aR[a][b] = f4m{0.f, 0.f, 0.f, 0.f}; aI[a][b] = f4m{0.f, 0.f, 0.f, 0.f}; }
    for (int I0 = 0; I0 < pb; I0 += KC) {
      for (int e = tid; e < TR * KC; e += 256) {
        const int r = e / KC, c = e % KC;
        const int row = r0 + r, I = I0 + c;
        float2 v = {0.f, 0.f};
        if (row < nrows && I < pb) { v = M[(size_t)row * pb + I]; v.x *= rs_[row]; v.y *= rs_[row]; }
        Ar[r][c] = v.x;
        Ai[r][c] = v.y;
      }
      for (int e = tid; e < KC * TJ; e += 256) {
        const int c = e % KC, j = e / KC;
        const int I = I0 + c, J = J0 + j;
        float2 v = {0.f, 0.f};
        if (I < pb && J < pb) {
          const double2 x = Ww[(size_t)I + (size_t)pb * J];
          v = float2{(float)(x.x * cs[J]), (float)(x.y * cs[J])};
        }
        Br[c][j] = v.x;
        Bi[c][j] = v.y;
      }
      __syncthreads();
      h8m arh[2], arl[2], aih[2], ail[2], anh[2], anl[2], brh[2], brl[2], bih[2], bil[2];
      for (int t = 0; t < 2; t++) {
        const int row = 32 * wr + 16 * t + li, col = 32 * wc + 16 * t + li;
        for (int j = 0; j < 8; j++) {
          const int kk = 8 * lk + j;
          _Float16 h, l;
          mw_split(Ar[row][kk], h, l); arh[t][j] = h; arl[t][j] = l;
          mw_split(Ai[row][kk], h, l); aih[t][j] = h; ail[t][j] = l; anh[t][j] = -h; anl[t][j] = -l;
          mw_split(Br[kk][col], h, l); brh[t][j] = h; brl[t][j] = l;
          mw_split(Bi[kk][col], h, l); bih[t][j] = h; bil[t][j] = l;
        }
      }
      // conj(m) W: Re = mr Wr + mi Wi, Im = mr Wi - mi Wr
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
          f4m r = aR[a][b], m = aI[a][b];
          r = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh[a], brh[b], r, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh[a], brl[b], r, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_16x16x32_f16(arl[a], brh[b], r, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih[a], bih[b], r, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih[a], bil[b], r, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_16x16x32_f16(ail[a], bih[b], r, 0, 0, 0);
          m = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh[a], bih[b], m, 0, 0, 0);
          m = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh[a], bil[b], m, 0, 0, 0);
          m = __builtin_amdgcn_mfma_f32_16x16x32_f16(arl[a], bih[b], m, 0, 0, 0);
          m = __builtin_amdgcn_mfma_f32_16x16x32_f16(anh[a], brh[b], m, 0, 0, 0);
          m = __builtin_amdgcn_mfma_f32_16x16x32_f16(anh[a], brl[b], m, 0, 0, 0);
          m = __builtin_amdgcn_mfma_f32_16x16x32_f16(anl[a], brh[b], m, 0, 0, 0);
          aR[a][b] = r;
          aI[a][b] = m;
        }
      __syncthreads();
    }
    for (int a = 0; a < 2; a++)
      for (int v = 0; v < 4; v++) {
        const int row = r0 + 32 * wr + 16 * a + 4 * lk + v;
        double2 s = rs[a][v];
        for (int b = 0; b < 2; b++) {
          const int J = J0 + 32 * wc + 16 * b + li;
          if (row < nrows && J < pb) {
            const double f = 1.0 / ((double)rs_[row] * (double)cs[J]);
            const float2 m = M[(size_t)row * pb + J];
            const double xr = (double)aR[a][b][v] * f, xi = (double)aI[a][b][v] * f;
            s.x += xr * (double)m.x - xi * (double)m.y;
            s.y += xr * (double)m.y + xi * (double)m.x;
          }
        }
        rs[a][v] = s;
      }
  }
  for (int a = 0; a < 2; a++)
    for (int v = 0; v < 4; v++) {
      double x = rs[a][v].x, y = rs[a][v].y;
      for (int off = 1; off < 16; off <<= 1) {
        x += __shfl_xor(x, off, 64);
        y += __shfl_xor(y, off, 64);
      }
      if (li == 0) racc[wc][32 * wr + 16 * a + 4 * lk + v] = double2{x, y};
    }
  __syncthreads();
  for (int r = tid; r < TR; r += 256) {
    const int row = r0 + r;
    if (row < nrows) {
      const double2 s0 = racc[0][r], s1 = racc[1][r];
      const int k = row / rpk, rr = row % rpk;
      out[((size_t)k * nom + w) * rpk + rr] = double2{wkq * (s0.x + s1.x), wkq * (s0.y + s1.y)};
    }
  }
}

// alpha_r = 2^(14-e), e = exponent of max_I max(|Re|,|Im|) of row r (complex64 M)
__global__ void pygw_mwm_rowscale(const float2* __restrict__ M, int nrows, int pb, float* __restrict__ rs) {
  const int row = blockIdx.x;
  float m = 0.f;
  for (int I = threadIdx.x; I < pb; I += blockDim.x) {
    const float2 v = M[(size_t)row * pb + I];
    m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
  }
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  __shared__ float red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < (int)(blockDim.x >> 6); i++) m = fmaxf(m, red[i]);
    m = fmaxf(m, red[0]);
    int e = 0;
    if (m > 0.f) frexpf(m, &e);
    rs[row] = ldexpf(1.0f, 14 - e);
  }
}

// beta_{J,w} = 2^(14-e), e = exponent of max_I max(|Re|,|Im|) of column J of W_w
__global__ void pygw_mwm_colscale(const double2* __restrict__ W, int pb, float* __restrict__ cs) {
  const int J = blockIdx.x, w = blockIdx.y;
  const double2* col = W + (size_t)w * pb * pb + (size_t)pb * J;
  double m = 0.0;
  for (int I = threadIdx.x; I < pb; I += blockDim.x) m = fmax(m, fmax(fabs(col[I].x), fabs(col[I].y)));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if (threadIdx.x == 0) {
    int e = 0;
    if (m > 0.0) frexp(m, &e);
    cs[(size_t)w * pb + J] = (float)ldexp(1.0, 14 - e);
  }
}

size_t mwm_mixed_ws_bytes(int nrows, int pb, int nom) {
  return sizeof(float) * ((size_t)nrows + (size_t)pb * nom) + 256;
}

int mwm_mixed_launch(const void* M, int nrows, int pb, const void* W, int nom, double wkq, int rpk, void* ws,
                     void* out, hipStream_t stream) {
  float* rs = (float*)ws;
  float* cs = rs + nrows;
  hipLaunchKernelGGL(pygw_mwm_rowscale, dim3(nrows), dim3(256), 0, stream, (const float2*)M, nrows, pb, rs);
  hipLaunchKernelGGL(pygw_mwm_colscale, dim3(pb, nom), dim3(64), 0, stream, (const double2*)W, pb, cs);
  dim3 grid(ceil_div(nrows, 64), nom);
  hipLaunchKernelGGL(pygw_mwm_mixed, grid, dim3(256), 0, stream, (const float2*)M, nrows, pb, (const double2*)W, nom,
                     wkq, rpk, rs, cs, (double2*)out);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}
