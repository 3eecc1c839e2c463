// Interstitial part of the product-basis matrix elements on gfx950
// (calc_minm_IS, for_vxcnn.f90:1030-1134, called at pygw.py:2661):
//
//   mmat[ie1][ie2][G] = 1/sqrt(Vol) sum_{i < nvi, j < nvj}
//       Aeigq[mstart+ie2][j] mpwipw[G][K(i, j)] Aeigk[nstart+ie1][i]
//
// with K(i, j) the product-basis plane wave equal to G_i - G_j + umklap
// (G_i of k, G_j of k - q; no term where there is none).
//
//   pygw_is_map    K(i, j) from the integer G vectors (i_g0, indgq_inverse)
//   pygw_is_gemm1  T[G][j][ie1] = sum_i mpwipw[G][K(i, j)] Aeigk[nstart+ie1][i]
//                  (the mpwipw row of G staged in LDS and gathered through K)
//   or, with workspace for P (the default): pygw_is_scatter puts
//                  P[K][j][ie1] = Aeigk[nstart+ie1][i] at K = K(i, j) (one i
//                  per (K, j): a k-point's plane waves are distinct), and
//                  pygw_is_gemm1p forms the same T as the dense GEMM
//                  T[G][(j, ie1)] = sum_K mpwipw[G][K] P[K][(j, ie1)]
//   pygw_is_gemm2  mmat[ie1][ie2][col0 + G] = sum_j Aeigq[mstart+ie2][j] T[G][j][ie1] / sqrt(Vol)
//
// Both products run on v_mfma_f64_16x16x4f64 (tiles of four waves of 32 x 32,
// 64 x 64 or 128 x 32 for a band window of at most 32; four real products
// per complex product).
#include "gemm_tile.h"

namespace is {
constexpr int kMaxBarc = 4096;  // mpwipw row staged in LDS (64 KB at most)
}


// kmap[i][j] = K(i, j) (0-based index into mpwipw's columns) or -1
// (for_vxcnn.f90:1083-1091: G = gindex[indgk_i[i]] - gindex[indgk_j[j]] +
// umklap; inside the i_g0 box, i_g0 gives its index in gindex, and
// indgq_inverse its index in the product basis)
// (indices outside gindex, i_g0 or the product basis count as none)
__global__ void __launch_bounds__(256) pygw_is_map(const int* __restrict__ gindex, int ng,
                                                  const int* __restrict__ indgk_i, const int* __restrict__ indgk_j,
                                                  int nvi, int nvj, const int* __restrict__ i_g0, int n1, int n2,
                                                  int n3, const int* __restrict__ inv, int inv_size, int ngq_barc,
                                                  int u0, int u1, int u2, int* __restrict__ kmap) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)nvi * nvj) return;
  const int i = (int)(e / nvj), j = (int)(e % nvj);
  const int a = indgk_i[i], b = indgk_j[j];
  if (a < 0 || a >= ng || b < 0 || b >= ng) {
    kmap[e] = -1;
    return;
  }
  const int* gi = gindex + 3 * (size_t)a;
  const int* gj = gindex + 3 * (size_t)b;
  const int x = gi[0] - gj[0] + u0, y = gi[1] - gj[1] + u1, z = gi[2] - gj[2] + u2;
  int K = -1;
  if (abs(x) <= n1 && abs(y) <= n2 && abs(z) <= n3) {
    const int idg = i_g0[((size_t)(x + n1) * (2 * n2 + 1) + (y + n2)) * (2 * n3 + 1) + (z + n3)];
    if (idg >= 0 && idg < inv_size) K = inv[idg];
    if (K >= ngq_barc) K = -1;
  }
  kmap[e] = K;
}

// T[G][j][ie1] = sum_i mpwipw[G][kmap[i][j]] Ak[nstart+ie1][i]; block: one
// G (blockIdx.y), TM j x TN ie1
template <int WR>
__global__ void __launch_bounds__(256) pygw_is_gemm1(const double2* __restrict__ mpwipw, int ngq_barc,
                                                    const int* __restrict__ kmap, int nvi, int nvj,
                                                    const double2* __restrict__ Ak, int ldk, int nstart, int ndim,
                                                    double2* __restrict__ T) {
  using namespace is;
  constexpr int TM = L<WR>::TM, TN = L<WR>::TN;
  extern __shared__ double2 row[];  // ngq_barc entries (dynamic: small rows leave room for more blocks)
  __shared__ double2 As[TM][KC + 1];
  __shared__ double2 Bs[KC][TN + 1];
  const int G = blockIdx.y;
  const int ntn = (ndim + TN - 1) / TN;
  const int j0 = (blockIdx.x / ntn) * TM, e0 = (blockIdx.x % ntn) * TN;
  for (int c = threadIdx.x; c < ngq_barc; c += 256) row[c] = mpwipw[(size_t)G * ngq_barc + c];
  __syncthreads();
  d4 aR[2][2], aI[2][2];
  zero(aR, aI);
  auto la = [&](int m, int k) -> double2 {  // A(j, i)
    const int j = j0 + m;
    if (j >= nvj || k >= nvi) return double2{0.0, 0.0};
    const int K = kmap[(size_t)k * nvj + j];
    return K >= 0 ? row[K] : double2{0.0, 0.0};
  };
  auto lb = [&](int k, int n) -> double2 {  // B(i, ie1)
    const int ie1 = e0 + n;
    return (k < nvi && ie1 < ndim) ? Ak[(size_t)(nstart + ie1) * ldk + k] : double2{0.0, 0.0};
  };
  tile<WR>(nvi, la, lb, As, Bs, aR, aI);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave / L<WR>::WC, wc = wave % L<WR>::WC, li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int j = j0 + 32 * wr + 16 * a + 4 * v + lk, ie1 = e0 + 32 * wc + 16 * b + li;
        if (j < nvj && ie1 < ndim) T[((size_t)G * nvj + j) * ndim + ie1] = double2{aR[a][b][v], aI[a][b][v]};
      }
}

// mmat[(ie1 mdim + ie2) ldo + col0 + G] = s sum_j Aq[mstart+ie2][j] T[G][j][ie1];
// block: one G (blockIdx.y), TM ie2 x TN ie1
template <int WR>
__global__ void __launch_bounds__(256) pygw_is_gemm2(const double2* __restrict__ Aq, int ldq, int mstart, int mdim,
                                                    const double2* __restrict__ T, int nvj, int ndim, double s,
                                                    double2* __restrict__ out, long ldo, int col0) {
  using namespace is;
  constexpr int TM = L<WR>::TM, TN = L<WR>::TN;
  __shared__ double2 As[TM][KC + 1];
  __shared__ double2 Bs[KC][TN + 1];
  const int G = blockIdx.y;
  const int ntn = (ndim + TN - 1) / TN;
  const int m0 = (blockIdx.x / ntn) * TM, e0 = (blockIdx.x % ntn) * TN;
  d4 aR[2][2], aI[2][2];
  zero(aR, aI);
  const double2* TG = T + (size_t)G * nvj * ndim;
  auto la = [&](int m, int k) -> double2 {  // A(ie2, j)
    const int ie2 = m0 + m;
    return (ie2 < mdim && k < nvj) ? Aq[(size_t)(mstart + ie2) * ldq + k] : double2{0.0, 0.0};
  };
  auto lb = [&](int k, int n) -> double2 {  // B(j, ie1)
    const int ie1 = e0 + n;
    return (k < nvj && ie1 < ndim) ? TG[(size_t)k * ndim + ie1] : double2{0.0, 0.0};
  };
  tile<WR>(nvj, la, lb, As, Bs, aR, aI);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave / L<WR>::WC, wc = wave % L<WR>::WC, li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int ie2 = m0 + 32 * wr + 16 * a + 4 * v + lk, ie1 = e0 + 32 * wc + 16 * b + li;
        if (ie2 < mdim && ie1 < ndim)
          out[((size_t)ie1 * mdim + ie2) * ldo + col0 + G] = double2{s * aR[a][b][v], s * aI[a][b][v]};
      }
}

// P[K][j][ie1] += Aeigk[nstart+ie1][i] for every (i, j) with K = K(i, j) >= 0
// (P zeroed first; one contribution per slot unless indgk_i repeats a
// plane wave, which the atomic sum then adds like the reference's sum)
__global__ void __launch_bounds__(256) pygw_is_scatter(const int* __restrict__ kmap, int nvi, int nvj,
                                                      const double2* __restrict__ Ak, int ldk, int nstart, int ndim,
                                                      double2* __restrict__ P) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)nvi * nvj * ndim) return;
  const int ie1 = (int)(e % ndim);
  const long ij = e / ndim;
  const int i = (int)(ij / nvj), j = (int)(ij % nvj);
  const int K = kmap[(size_t)i * nvj + j];
  if (K < 0) return;
  const double2 a = Ak[(size_t)(nstart + ie1) * ldk + i];
  double* d = (double*)(P + ((size_t)K * nvj + j) * ndim + ie1);
  atomicAdd(d, a.x);
  atomicAdd(d + 1, a.y);
}

__global__ void __launch_bounds__(256) pygw_is_zero(double2* __restrict__ P, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    P[i] = double2{0.0, 0.0};
}

// T[G][n] = sum_K mpwipw[G][K] P[K][n] (n = j ndim + ie1 < N): a dense GEMM.
// Items (column tile, row tile), row tiles innermost, dealt to the XCDs in
// contiguous runs so the row tiles of a column slice of P share an L2.
__global__ void __launch_bounds__(256) pygw_is_gemm1p(const double2* __restrict__ mpwipw, int ngq, int ngq_barc,
                                                     const double2* __restrict__ P, long N, long per,
                                                     double2* __restrict__ T) {
  using namespace is;
  __shared__ double2 As[64][KC + 1];
  __shared__ double2 Bs[KC][64 + 1];
  const int ntm = (ngq + 63) / 64;
  const long item = (long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (item >= ((N + 63) / 64) * ntm) return;
  const int m0 = (int)(item % ntm) * 64;
  const long n0 = item / ntm * 64;
  d4 aR[2][2], aI[2][2];
  zero(aR, aI);
  auto la = [&](int m, int k) -> double2 {
    const int r = m0 + m;
    return (r < ngq && k < ngq_barc) ? mpwipw[(size_t)r * ngq_barc + k] : double2{0.0, 0.0};
  };
  auto lb = [&](int k, int c) -> double2 {
    const long col = n0 + c;
    return (k < ngq_barc && col < N) ? P[(size_t)k * N + col] : double2{0.0, 0.0};
  };
  tile<2>(ngq_barc, la, lb, As, Bs, aR, aI);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave >> 1, wc = wave & 1, li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = m0 + 32 * wr + 16 * a + 4 * v + lk;
        const long col = n0 + 32 * wc + 16 * b + li;
        if (r < ngq && col < N) T[(size_t)r * N + col] = double2{aR[a][b][v], aI[a][b][v]};
      }
}

size_t minm_is_ws_bytes(int ngq, int nvi, int nvj, int ndim) {
  (void)ngq;
  return (((size_t)nvi * nvj * 4 + 255) & ~(size_t)255) + (size_t)ngq * nvj * ndim * 16;
}

// the dense form also holds P [ngq_barc][nvj][ndim]
size_t minm_is_ws_bytes_dense(int ngq, int ngq_barc, int nvi, int nvj, int ndim) {
  return minm_is_ws_bytes(ngq, nvi, nvj, ndim) + (size_t)ngq_barc * nvj * ndim * 16;
}

int minm_is_launch(int nstart, int nend, int mstart, int mend, const void* Aeigk, int ldk, const void* Aeigq,
                   int ldq, const int* umk, const void* mpwipw, int ngq, int ngq_barc, int nvi, int nvj,
                   const int* indgk_i, const int* indgk_j, const int* inv, int inv_size, const int* gindex,
                   int ng, const int* i_g0, int n1, int n2, int n3, double vol, void* mmat, long ldo, int col0, void* ws,
                   size_t ws_size, hipStream_t stream) {
  using namespace is;
  const int ndim = nend - nstart, mdim = mend - mstart;
  // with room for P, the first product is the dense GEMM over the product
  // basis; otherwise the gather form, whose LDS row limits ngq_barc
  const bool dense = ws_size >= minm_is_ws_bytes_dense(ngq, ngq_barc, nvi, nvj, ndim);
  if (ws_size < minm_is_ws_bytes(ngq, nvi, nvj, ndim) || (!dense && ngq_barc > kMaxBarc)) return PYGW_E_ARG;
  int* kmap = (int*)ws;
  double2* T = (double2*)((char*)ws + (((size_t)nvi * nvj * 4 + 255) & ~(size_t)255));
  hipLaunchKernelGGL(pygw_is_map, dim3(ceil_div((long)nvi * nvj, 256)), dim3(256), 0, stream, gindex, ng,
                     indgk_i, indgk_j, nvi, nvj, i_g0, n1, n2, n3, inv, inv_size, ngq_barc, umk[0], umk[1], umk[2],
                     kmap);
  auto run = [&](auto wr) {
    constexpr int WR = decltype(wr)::value, TM = L<WR>::TM, TN = L<WR>::TN;
    const int ntn = ceil_div(ndim, TN);
    if (dense) {
      double2* P = T + (size_t)ngq * nvj * ndim;
      const long nP = (long)ngq_barc * nvj * ndim, N = (long)nvj * ndim;
      hipLaunchKernelGGL(pygw_is_zero, dim3((unsigned)std::min<long>(ceil_div(nP, 256), 1L << 20)), dim3(256), 0,
                         stream, P, nP);
      hipLaunchKernelGGL(pygw_is_scatter, dim3(ceil_div((long)nvi * nvj * ndim, 256)), dim3(256), 0, stream, kmap,
                         nvi, nvj, (const double2*)Aeigk, ldk, nstart, ndim, P);
      const long items = (N + 63) / 64 * ceil_div(ngq, 64), per = (items + 7) / 8;
      hipLaunchKernelGGL(pygw_is_gemm1p, dim3((unsigned)(8 * per)), dim3(256), 0, stream, (const double2*)mpwipw,
                         ngq, ngq_barc, P, N, per, T);
    } else {
      hipLaunchKernelGGL(pygw_is_gemm1<WR>, dim3(ceil_div(nvj, TM) * ntn, ngq), dim3(256), ngq_barc * 16, stream,
                         (const double2*)mpwipw, ngq_barc, kmap, nvi, nvj, (const double2*)Aeigk, ldk, nstart, ndim,
                         T);
    }
    hipLaunchKernelGGL(pygw_is_gemm2<WR>, dim3(ceil_div(mdim, TM) * ntn, ngq), dim3(256), 0, stream,
                       (const double2*)Aeigq, ldq, mstart, mdim, T, nvj, ndim, 1.0 / sqrt(vol), (double2*)mmat, ldo,
                       col0);
  };
  // a band window of at most 32 (the occupied bands of the polarization)
  // runs on 128 x 32 tiles, wider ones on 64 x 64
  if (ndim <= 32)
    run(std::integral_constant<int, 4>());
  else
    run(std::integral_constant<int, 2>());
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}
