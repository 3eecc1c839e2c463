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
// or, when it costs fewer products and the workspace holds it, the pair
// density first (the same sum, associated the other way):
//   pygw_is_imap_fill, pygw_is_imap  imap[K][j] = the i with K(i, j) = K
//                        (-1: none; -2: several, when k repeats a plane wave)
//   pygw_is_gemm_ab      rho[ie2][(ie1, K)] = sum_j Aeigq[mstart+ie2][j] Q(j, (ie1, K)),
//                        Q(j, (ie1, K)) = Aeigk[nstart+ie1][imap[K][j]] gathered
//                        while the tile is loaded (the sum over those i for
//                        -2), split over j in fixed slices that
//                        pygw_is_reduce adds in order into rho[(ie1, ie2)][K]
//   pygw_is_gemm_ab      mmat[(ie1, ie2)][col0 + G] = sum_K rho[(ie1, ie2)][K] mpwipw[G][K] / sqrt(Vol)
//                        (split over K the same way)
// 8 ndim ngq_barc mdim (nvj + ngq) FLOP against 8 ngq ndim nvj (ngq_barc +
// mdim): at the polarization's 20 x 120 window, 600 plane waves and 400 of
// 700 basis plane waves, 1.34e10 against 3.15e10 per k.
//
// Both products run on v_mfma_f64_16x16x4f64 (tiles of four waves of 32 x 32,
// 64 x 64 or 128 x 32 for a band window of at most 32; four real products
// per complex product).
#include <algorithm>

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

// imap[K][j] = -1 (nvj ints per product-basis plane wave)
__global__ void __launch_bounds__(256) pygw_is_imap_fill(int* __restrict__ imap, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) imap[e] = -1;
}

// imap[K][j] = i for K = K(i, j) >= 0: a k-point's plane waves are distinct,
// so one i per (K, j); a second one marks the slot -2 (the gather then sums
// every such i, as the reference's sum over i does)
__global__ void __launch_bounds__(256) pygw_is_imap(const int* __restrict__ kmap, int nvi, int nvj,
                                                   int* __restrict__ imap) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)nvi * nvj) return;
  const int K = kmap[e];
  if (K < 0) return;
  const int i = (int)(e / nvj), j = (int)(e % nvj);
  int* slot = imap + (size_t)K * nvj + j;
  if (atomicCAS(slot, -1, i) != -1) atomicExch(slot, -2);
}

namespace is {
// B(k, n) of the first pair-density product: Q(j, (ie1, K)) gathered from
// Aeigk through imap (n = ie1 nb + K)
struct GatherB {
  const int* imap;
  const int* kmap;
  const double2* Ak;
  int ldk, nstart, nvi, nvj, nb;
  __device__ __forceinline__ double2 operator()(int j, long n) const {
    const int ie1 = (int)(n / nb), K = (int)(n - (long)ie1 * nb);
    const int v = imap[(size_t)K * nvj + j];
    const double2* row = Ak + (size_t)(nstart + ie1) * ldk;
    if (v >= 0) return row[v];
    double2 a{0.0, 0.0};
    if (v == -2)
      for (int i = 0; i < nvi; i++)
        if (kmap[(size_t)i * nvj + j] == K) {
          a.x += row[i].x;
          a.y += row[i].y;
        }
    return a;
  }
};
}  // namespace is

// W[s][m][n] = sum_{k in slice s} A[m lda + k] B(k, n), with B(k, n) =
// B[k ldb + n] (BM 0), B[n ldb + k] (BM 1) or gathered by g (BM 2); slice s
// covers k in [s kc, min(K, (s + 1) kc)).  Items (column tile, slice, row
// tile) with row tiles innermost (rows_inner: the row tiles of one column
// slice of B run together), else (row tile, slice, column tile), dealt to
// the XCDs in contiguous runs so that the tiles sharing an operand slice
// share an L2; 64 x 64 tiles.
template <int BM>
__global__ void __launch_bounds__(256) pygw_is_gemm_ab(const double2* __restrict__ A, long lda,
                                                      const double2* __restrict__ B, long ldb, is::GatherB g, int M,
                                                      long N, int K, int kc, int splits, int rows_inner, long per,
                                                      double2* __restrict__ W) {
  using namespace is;
  __shared__ double2 As[64][KC + 1];
  __shared__ double2 Bs[KC][64 + 1];
  const int ntm = (M + 63) / 64;
  const long ntn = (N + 63) / 64;
  const long item = (long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (item >= ntn * ntm * splits) return;
  int m0, s;
  long n0;
  if (rows_inner) {
    m0 = (int)(item % ntm) * 64;
    const long rest = item / ntm;
    s = (int)(rest % splits);
    n0 = rest / splits * 64;
  } else {
    n0 = item % ntn * 64;
    const long rest = item / ntn;
    s = (int)(rest % splits);
    m0 = (int)(rest / splits) * 64;
  }
  const int k0 = s * kc, kn = min(K - k0, kc);
  d4 aR[2][2], aI[2][2];
  zero(aR, aI);
  auto la = [&](int m, int k) -> double2 {
    const int r = m0 + m;
    return (r < M && k < kn) ? A[(size_t)r * lda + k0 + k] : double2{0.0, 0.0};
  };
  auto lb = [&](int k, int c) -> double2 {
    const long col = n0 + c;
    if (k >= kn || col >= N) return double2{0.0, 0.0};
    if (BM == 2) return g(k0 + k, col);
    return BM == 1 ? B[(size_t)col * ldb + k0 + k] : B[(size_t)(k0 + k) * ldb + col];
  };
  tile<2>(kn, la, lb, As, Bs, aR, aI);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave >> 1, wc = wave & 1, li = lane & 15, lk = lane >> 4;
  double2* Ws = W + (size_t)s * M * N;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = m0 + 32 * wr + 16 * a + 4 * v + lk;
        const long col = n0 + 32 * wc + 16 * b + li;
        if (r < M && col < N) Ws[(size_t)r * N + col] = double2{aR[a][b][v], aI[a][b][v]};
      }
}

// out = scale sum_s W[s][m][n], the slices added in order: nb = 0 puts (m, n)
// at out[m ldo + col0 + n]; nb > 0 reads n as (n / nb, n % nb) and puts it
// at out[((n / nb) M + m) ldo + col0 + n % nb]
__global__ void __launch_bounds__(256) pygw_is_reduce(const double2* __restrict__ W, int splits, int M, long N,
                                                     double scale, int nb, double2* __restrict__ out, long ldo,
                                                     int col0) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)M * N) return;
  const int m = (int)(e / N);
  const long n = e % N;
  double2 a = W[e];
  for (int s = 1; s < splits; s++) {
    const double2 b = W[(size_t)s * M * N + e];
    a.x += b.x;
    a.y += b.y;
  }
  const size_t o = nb > 0 ? ((size_t)(n / nb) * M + m) * ldo + col0 + n % nb : (size_t)m * ldo + col0 + n;
  out[o] = double2{scale * a.x, scale * a.y};
}

namespace is {
// slices of the inner dimension so that tiles x slices reaches ~1024 blocks,
// no slice shorter than 128 (a multiple of the 16-wide chunk)
struct Split {
  int kc, splits;
};
inline Split split_k(long tiles, int K) {
  long s = std::max<long>(1, std::min<long>((1024 + tiles - 1) / tiles, K / 128));
  int kc = (int)(((K + s - 1) / s + KC - 1) / KC * KC);
  return Split{kc, (K + kc - 1) / kc};
}
inline long tiles64(long M, long N) { return ((M + 63) / 64) * ((N + 63) / 64); }

struct RhoLayout {
  size_t kmap, imap, rho, W, total;
  Split s1, s2;
};
inline RhoLayout rho_layout(int ngq, int ngq_barc, int nvi, int nvj, int ndim, int mdim) {
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  RhoLayout r;
  const long n1 = (long)ndim * ngq_barc, m2 = (long)ndim * mdim;
  r.s1 = split_k(tiles64(mdim, n1), nvj);
  r.s2 = split_k(tiles64(m2, ngq), ngq_barc);
  r.kmap = 0;
  r.imap = al((size_t)nvi * nvj * 4);
  r.rho = r.imap + al((size_t)ngq_barc * nvj * 4);
  r.W = r.rho + al((size_t)m2 * ngq_barc * 16);
  r.total = r.W + std::max((size_t)r.s1.splits * mdim * n1, (size_t)r.s2.splits * m2 * ngq) * 16;
  return r;
}
// FLOP of the two associations (complex multiply-adds x 8)
inline bool rho_cheaper(int ngq, int ngq_barc, int nvj, int ndim, int mdim) {
  const double rho = (double)ndim * ngq_barc * mdim * ((double)nvj + ngq);
  const double t = (double)ngq * ndim * nvj * ((double)ngq_barc + mdim);
  return rho < t;
}
}  // namespace is

size_t minm_is_ws_bytes_rho(int ngq, int ngq_barc, int nvi, int nvj, int ndim, int mdim) {
  return is::rho_layout(ngq, ngq_barc, nvi, nvj, ndim, mdim).total;
}

bool minm_is_rho_cheaper(int ngq, int ngq_barc, int nvj, int ndim, int mdim) {
  return is::rho_cheaper(ngq, ngq_barc, nvj, ndim, mdim);
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
  // the pair-density association when it costs fewer products and the
  // workspace holds it; else, with room for P, the first product as the
  // dense GEMM over the product basis; else the gather form, whose LDS row
  // limits ngq_barc
  const bool rho = rho_cheaper(ngq, ngq_barc, nvj, ndim, mdim) &&
                   ws_size >= minm_is_ws_bytes_rho(ngq, ngq_barc, nvi, nvj, ndim, mdim);
  const bool dense = !rho && ws_size >= minm_is_ws_bytes_dense(ngq, ngq_barc, nvi, nvj, ndim);
  if (!rho && (ws_size < minm_is_ws_bytes(ngq, nvi, nvj, ndim) || (!dense && ngq_barc > kMaxBarc)))
    return PYGW_E_ARG;
  int* kmap = (int*)ws;
  if (rho) {
    const RhoLayout r = rho_layout(ngq, ngq_barc, nvi, nvj, ndim, mdim);
    int* imap = (int*)((char*)ws + r.imap);
    double2* R = (double2*)((char*)ws + r.rho);
    double2* W = (double2*)((char*)ws + r.W);
    hipLaunchKernelGGL(pygw_is_map, dim3(ceil_div((long)nvi * nvj, 256)), dim3(256), 0, stream, gindex, ng,
                       indgk_i, indgk_j, nvi, nvj, i_g0, n1, n2, n3, inv, inv_size, ngq_barc, umk[0], umk[1], umk[2],
                       kmap);
    const long nI = (long)ngq_barc * nvj, N1 = (long)ndim * ngq_barc, M2 = (long)ndim * mdim;
    hipLaunchKernelGGL(pygw_is_imap_fill, dim3((unsigned)std::min<long>(ceil_div(nI, 256), 1L << 16)), dim3(256), 0,
                       stream, imap, nI);
    hipLaunchKernelGGL(pygw_is_imap, dim3(ceil_div((long)nvi * nvj, 256)), dim3(256), 0, stream, kmap, nvi, nvj,
                       imap);
    const GatherB g{imap, kmap, (const double2*)Aeigk, ldk, nstart, nvi, nvj, ngq_barc};
    long items = tiles64(mdim, N1) * r.s1.splits, per = (items + 7) / 8;
    hipLaunchKernelGGL(pygw_is_gemm_ab<2>, dim3((unsigned)(8 * per)), dim3(256), 0, stream,
                       (const double2*)Aeigq + (size_t)mstart * ldq, (long)ldq, (const double2*)nullptr, 0L, g, mdim,
                       N1, nvj, r.s1.kc, r.s1.splits, 1, per, W);
    hipLaunchKernelGGL(pygw_is_reduce, dim3(ceil_div((long)mdim * N1, 256)), dim3(256), 0, stream, W, r.s1.splits,
                       mdim, N1, 1.0, ngq_barc, R, (long)ngq_barc, 0);
    items = tiles64(M2, ngq) * r.s2.splits;
    per = (items + 7) / 8;
    hipLaunchKernelGGL(pygw_is_gemm_ab<1>, dim3((unsigned)(8 * per)), dim3(256), 0, stream, (const double2*)R,
                       (long)ngq_barc, (const double2*)mpwipw, (long)ngq_barc, g, (int)M2, (long)ngq, ngq_barc,
                       r.s2.kc, r.s2.splits, 0, per, W);
    hipLaunchKernelGGL(pygw_is_reduce, dim3(ceil_div(M2 * ngq, 256)), dim3(256), 0, stream, W, r.s2.splits, (int)M2,
                       (long)ngq, 1.0 / sqrt(vol), 0, (double2*)mmat, ldo, col0);
    return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
  }
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
