// Complex FP64 GEMM tile on v_mfma_f64_16x16x4f64 with operand loaders,
// shared by the interstitial matrix elements and the plane-wave basis.
#pragma once
#include "common.h"

typedef double d4 __attribute__((ext_vector_type(4)));

namespace is {
constexpr int KC = 16;

// Tile of WR x (4 / WR) waves, each 32 x 32 (2 x 2 MFMA tiles): 64 x 64 (WR
// = 2) or 128 x 32 (WR = 4, for a narrow band window)
template <int WR>
struct L {
  static constexpr int WC = 4 / WR, TM = 32 * WR, TN = 32 * WC;
};

// C[m][n] += sum_k A(m, k) B(k, n) over one tile: la(m, k) and lb(k, n)
// give the operands (zero outside); acc rows 32 wr + 16 a + 4 v + lk,
// columns 32 wc + 16 b + li
template <int WR, class LA, class LB>
__device__ __forceinline__ void tile(int K, LA la, LB lb, double2 (&As)[L<WR>::TM][KC + 1],
                                     double2 (&Bs)[KC][L<WR>::TN + 1], d4 (&aR)[2][2], d4 (&aI)[2][2]) {
  constexpr int TM = L<WR>::TM, TN = L<WR>::TN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave / L<WR>::WC, wc = wave % L<WR>::WC, li = lane & 15, lk = lane >> 4;
  for (int k0 = 0; k0 < K; k0 += KC) {
    for (int e = tid; e < TM * KC; e += 256) {
      const int m = e % TM, c = e / TM;  // consecutive threads: consecutive m
      As[m][c] = la(m, k0 + c);
    }
    for (int e = tid; e < KC * TN; e += 256) {
      const int c = e % KC, n = e / KC;  // consecutive threads: consecutive k
      Bs[c][n] = lb(k0 + c, n);
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
        ar[t] = x.x;
        ai[t] = x.y;
        br[t] = y.x;
        bi[t] = y.y;
      }
      // a b: Re = ar br - ai bi, Im = ar bi + ai br
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
          aR[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[b], aR[a][b], 0, 0, 0);
          aR[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[a], bi[b], aR[a][b], 0, 0, 0);
          aI[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], bi[b], aI[a][b], 0, 0, 0);
          aI[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], br[b], aI[a][b], 0, 0, 0);
        }
    }
    __syncthreads();
  }
}

__device__ __forceinline__ void zero(d4 (&aR)[2][2], d4 (&aI)[2][2]) {
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      aR[a][b] = d4{0, 0, 0, 0};
      aI[a][b] = d4{0, 0, 0, 0};
    }
}
}  // namespace is
