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
