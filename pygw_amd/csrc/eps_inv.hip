// Batched dielectric inversion of calc_eps (pygw.py:3229-3271), FP64.
//
// Per Matsubara frequency:  A = (eps + eps^H)/2 + 1;  eps <- A^{-1} - 1.
// A is Hermitian positive definite (the weights are <= 0 and enter with
// coefficient -2, so eps is positive semidefinite), hence the inverse is
// formed by block Gauss-Jordan elimination without pivoting, NB columns at a
// time; every Schur complement of an HPD matrix is HPD, so each NB x NB
// pivot block is invertible.  One elimination step P:
//   B = A_PP^{-1};  A_ij -= A_iP B A_Pj;  A_iP <- -A_iP B;  A_Pj <- B A_Pj;
//   A_PP <- B.
// All frequencies advance together (grid.z = frequency), so each step is
// three launches over n_om matrices.  Matrices are column-major, contiguous
// per frequency (PyGW's eps[:, :, iom] of a Fortran [PB, PB, n_om] array).
//
// Only the 64 x 64 tiles on and above the diagonal are eliminated.  With s_i
// = -1 for the indices already swept and +1 otherwise, every step keeps
// A_ji = s_i s_j conj(A_ij) (a Hermitian A; the step negates the coupling of
// the pivot block to the unswept indices), so the strictly lower tiles follow
// from the upper ones: the panels of a step read them through that relation,
// and once every index is swept the inverse is Hermitian and the lower
// triangle is its mirror.  The update then costs about half of the full
// rank-nb update.
#include "common.h"

namespace ei {
constexpr int NB = 64;   // pivot block: each elimination step reads and writes A once
constexpr int KH = 32;   // k half staged at a time by the update
}

// (A + A^H)/2 + I, in place: one block per 32 x 32 tile pair (I <= J) of a
// frequency, both tiles staged in LDS so that every read and write walks
// consecutive rows of the column-major matrix.
__global__ void __launch_bounds__(256) pygw_ei_hermitize(double2* __restrict__ A, int n) {
  __shared__ double2 X[32][33], Y[32][33];
  const size_t w = blockIdx.z;
  double2* a = A + w * (size_t)n * n;
  const int I = blockIdx.x, J = blockIdx.y;
  if (I > J) return;
  const int i0 = 32 * I, j0 = 32 * J, tr = threadIdx.x & 31, tc = threadIdx.x >> 5;
  for (int c = tc; c < 32; c += 8) {
    const int i = i0 + tr, j = j0 + c, ii = j0 + tr, jj = i0 + c;
    X[c][tr] = (i < n && j < n) ? a[i + (size_t)n * j] : double2{0.0, 0.0};     // X[c][r] = A[i0 + r][j0 + c]
    Y[c][tr] = (ii < n && jj < n) ? a[ii + (size_t)n * jj] : double2{0.0, 0.0};  // Y[c][r] = A[j0 + r][i0 + c]
  }
  __syncthreads();
  for (int c = tc; c < 32; c += 8) {
    // h = (A[i][j] + conj(A[j][i]))/2 at i = i0 + tr, j = j0 + c; A[j][i] = Y[tr][c]
    const int i = i0 + tr, j = j0 + c;
    const double2 x = X[c][tr], y = Y[tr][c];
    double2 h{(x.x + y.x) / 2.0, (x.y - y.y) / 2.0};
    if (i == j) h = double2{h.x + 1.0, 0.0};
    if (i < n && j < n) a[i + (size_t)n * j] = h;
    // the mirror entry A[j0 + tr][i0 + c] = conj(h at i = i0 + c, j = j0 + tr)
    const int ii = j0 + tr, jj = i0 + c;
    const double2 x2 = X[tr][c], y2 = Y[c][tr];
    double2 h2{(x2.x + y2.x) / 2.0, (x2.y - y2.y) / 2.0};
    if (jj == ii) h2 = double2{h2.x + 1.0, 0.0};
    if (ii < n && jj < n) a[ii + (size_t)n * jj] = double2{h2.x, -h2.y};
  }
}

// A - I in place.
__global__ void pygw_ei_minus_identity(double2* __restrict__ A, int n, int nom) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * nom) return;
  const int w = i / n, d = i % n;
  double2* a = A + (size_t)w * n * n + d + (size_t)n * d;
  a->x = a->x - 1.0;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
  // Smith's algorithm
  if (fabs(b.x) >= fabs(b.y)) {
    double r = b.y / b.x, d = b.x + b.y * r;
    return double2{(a.x + a.y * r) / d, (a.y - a.x * r) / d};
  }
  double r = b.x / b.y, d = b.x * r + b.y;
  return double2{(a.x * r + a.y) / d, (a.y * r - a.x) / d};
}

// Step 1: invert the pivot block (size nb <= NB) of every frequency in LDS
// by unblocked Gauss-Jordan; B[w] is written column-major nb x nb.
// One block of 1024 threads per frequency (the frequencies alone fill a
// quarter of the CUs; each elimination step is 4 entries per thread between
// two barriers).  info[w] |= 1 if a pivot is not positive (matrix not
// positive definite).
__global__ void __launch_bounds__(1024) pygw_ei_pivot(const double2* __restrict__ A, int n, int p0, int nb,
                                                      double2* __restrict__ B, int* __restrict__ info) {
  using namespace ei;
  __shared__ double2 S[NB][NB + 1];
  const size_t w = blockIdx.x;
  const double2* a = A + w * (size_t)n * n;
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
    const int i = e % nb, j = e / nb;
    S[i][j] = a[(p0 + i) + (size_t)n * (p0 + j)];
  }
  __syncthreads();
  // Each thread owns the entries e = tid, tid + 1024, ... (at most 4): per
  // step it forms their new values from the pre-step row and column k
  // (reads), then, after one barrier, stores them (writes): two barriers per
  // step.
  constexpr int PER = (NB * NB + 1023) / 1024;
  for (int k = 0; k < nb; k++) {
    const double2 piv = S[k][k];
    const double2 rp = cdiv(double2{1.0, 0.0}, piv);
    if (threadIdx.x == 0 && !(piv.x > 0.0)) atomicOr(&info[w], 1);
    double2 nv[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int e = threadIdx.x + 1024 * q;
      if (e >= nb * nb) break;
      const int i = e % nb, j = e / nb;
      if (i != k && j != k) {
        const double2 t = cmul(S[i][k], cmul(S[k][j], rp));
        nv[q] = double2{S[i][j].x - t.x, S[i][j].y - t.y};
      } else if (j == k && i != k) {
        const double2 c = cmul(S[i][k], rp);
        nv[q] = double2{-c.x, -c.y};
      } else if (i == k && j != k) {
        nv[q] = cmul(S[k][j], rp);
      } else {
        nv[q] = rp;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int e = threadIdx.x + 1024 * q;
      if (e >= nb * nb) break;
      S[e % nb][e / nb] = nv[q];
    }
    __syncthreads();
  }
  double2* b = B + w * NB * NB;
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
    const int i = e % nb, j = e / nb;
    b[i + NB * j] = S[i][j];
  }
}

// Step 2: panels.  U[w] (nb x n, row panel) = B * A[P, :]  and
// L[w] (n x nb, column panel) = A[:, P] (copy before the update).
// grid (ceil(n/64), nom), 256 threads: 64 columns j per block, A[P, j] and B
// staged in LDS, each thread forms 16 rows of one column of U and copies
// 16 entries of L.
__global__ void __launch_bounds__(256) pygw_ei_panels(const double2* __restrict__ A, int n, int p0, int nb,
                                                      const double2* __restrict__ B, double2* __restrict__ U,
                                                      double2* __restrict__ L) {
  using namespace ei;
  typedef double d4v __attribute__((ext_vector_type(4)));
  __shared__ double Br[NB][NB + 1], Bi[NB][NB + 1];  // [c][r] = B[r][c]
  __shared__ double Cr[NB][64 + 1], Ci[NB][64 + 1];  // [c][jj] = A[p0 + c][j0 + jj]
  const size_t w = blockIdx.y;
  const double2* a = A + w * (size_t)n * n;
  const double2* b = B + w * NB * NB;
  const int j0 = blockIdx.x * 64;
  for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) {
    const int r = e % NB, c = e / NB;     // consecutive threads: consecutive rows of column c of B
    const double2 v = (r < nb && c < nb) ? b[r + NB * c] : double2{0.0, 0.0};
    Br[c][r] = v.x;
    Bi[c][r] = v.y;
  }
  auto put = [&](int c, int jj, double2 v) {
    Cr[c][jj] = v.x;
    Ci[c][jj] = v.y;
  };
  if (j0 + 64 <= p0) {
    // every j of the tile swept: A[P, j] = -conj(A[j, P]) (stored above the
    // diagonal), read down the columns p0 + c (consecutive threads:
    // consecutive rows j)
    for (int e = threadIdx.x; e < NB * 64; e += blockDim.x) {
      const int jj = e % 64, c = e / 64;
      double2 v{0.0, 0.0};
      if (c < nb) {
        const double2 t = a[(j0 + jj) + (size_t)n * (p0 + c)];
        v = double2{-t.x, t.y};
      }
      put(c, jj, v);
    }
  } else {
    for (int e = threadIdx.x; e < NB * 64; e += blockDim.x) {
      const int c = e % NB, jj = e / NB;  // consecutive threads: consecutive rows of one column
      const int j = j0 + jj;
      double2 v{0.0, 0.0};
      if (c < nb && j < n) {
        if (j >= p0) {
          v = a[(p0 + c) + (size_t)n * j];
        } else {  // A[P, j] of a swept j: -conj(A[j, P]) (stored above the diagonal)
          const double2 t = a[j + (size_t)n * (p0 + c)];
          v = double2{-t.x, t.y};
        }
      }
      put(c, jj, v);
    }
  }
  __syncthreads();
  // U[r][j] = sum_c B[r][c] A[P_c][j] on v_mfma_f64_16x16x4f64, formed as
  // D[j][r] (D row 4v + lane/16 indexes j, D column lane%16 indexes r), so
  // each group of 16 lanes writes 16 consecutive rows r of column j of U;
  // 4 waves of 32 (j) x 32 (r)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wj = wave & 1, wr = wave >> 1, li = lane & 15, lk = lane >> 4;
  d4v sr[2][2], si[2][2];  // [j tile][r tile]
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) { sr[x][y] = d4v{0, 0, 0, 0}; si[x][y] = d4v{0, 0, 0, 0}; }
#pragma unroll 4
  for (int ks = 0; ks < NB / 4; ks++) {
    const int c = 4 * ks + lk;
    double cr[2], ci[2], nci[2], br[2], bi[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      cr[t] = Cr[c][32 * wj + 16 * t + li];
      ci[t] = Ci[c][32 * wj + 16 * t + li];
      nci[t] = -ci[t];
      br[t] = Br[c][32 * wr + 16 * t + li];
      bi[t] = Bi[c][32 * wr + 16 * t + li];
    }
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
      for (int y = 0; y < 2; y++) {
        sr[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(cr[x], br[y], sr[x][y], 0, 0, 0);
        sr[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(nci[x], bi[y], sr[x][y], 0, 0, 0);
        si[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(cr[x], bi[y], si[x][y], 0, 0, 0);
        si[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ci[x], br[y], si[x][y], 0, 0, 0);
      }
  }
  double2* u = U + w * (size_t)NB * n;
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int j = j0 + 32 * wj + 16 * x + 4 * v + lk;
        const int r = 32 * wr + 16 * y + li;
        if (j < n && r < nb) u[r + (size_t)NB * j] = double2{sr[x][y][v], si[x][y][v]};
      }
  // L: A[j, P] (the rows of the pivot columns), below the pivot block from
  // the relation conj(A[P, j]) (both unswept)
  double2* l = L + w * (size_t)n * NB;
  const int jj = threadIdx.x & 63, rg = threadIdx.x >> 6;
  const int j = j0 + jj;
  if (j < n)
    for (int c = 16 * rg; c < 16 * rg + 16 && c < nb; c++)
      l[j + (size_t)n * c] = j < p0 + nb ? a[j + (size_t)n * (p0 + c)] : double2{Cr[c][jj], -Ci[c][jj]};
}

// Step 3: rank-nb update of every 64 x 64 tile.
//   rows i in P, cols j in P : A_PP = B
//   rows i in P              : A_Pj = U_j
//   cols j in P              : A_iP = -L_i B
//   otherwise                : A_ij -= L_i U_j
typedef double d4 __attribute__((ext_vector_type(4)));

// Step 3 on FP64 MFMA: 64 x 64 tiles, 4 waves of 32 x 32 (2 x 2 tiles of
// v_mfma_f64_16x16x4f64).  The products are formed transposed, D = U^T L^T
// (D row = 4v + lane/16 indexes j, D column = lane%16 indexes i), so each
// group of 16 lanes reads and writes 16 consecutive rows i of the
// column-major matrix.  S = L U in complex: Re = Lr Ur - Li Ui,
// Im = Lr Ui + Li Ur (four real products per k step).  Same cases as
// pygw_ei_update.
__global__ void __launch_bounds__(256) pygw_ei_update_mfma(double2* __restrict__ A, int n, int p0, int nb,
                                                           const double2* __restrict__ B,
                                                           const double2* __restrict__ U,
                                                           const double2* __restrict__ L, int nom, int ntu,
                                                           long per) {
  using namespace ei;
  __shared__ double Lr[KH][64 + 1], Li[KH][64 + 1];  // [k - k0][i]
  __shared__ double Ur[KH][64 + 1], Ui[KH][64 + 1];  // [k - k0][j]
  // work items (frequency, upper tile), frequency-major, dealt to the XCDs in
  // contiguous runs: the blocks of one frequency run on one XCD one after
  // another, so its panels U and L (1.8 MB at n = 900) stay in that XCD's L2
  // while its ~120 tiles read them
  const long item = (long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (item >= (long)nom * ntu) return;
  const size_t w = (size_t)(item / ntu);
  const int t = (int)(item % ntu);
  // upper tile t -> (ti <= tj), the tiles of one column tile together
  int tj = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((tj + 1) * (tj + 2) / 2 <= t) tj++;
  while (tj * (tj + 1) / 2 > t) tj--;
  const int i0 = (t - tj * (tj + 1) / 2) * 64, j0 = tj * 64;
  double2* a = A + w * (size_t)n * n;
  const double2* u = U + w * (size_t)NB * n;
  const double2* l = L + w * (size_t)n * NB;
  const double2* b = B + w * NB * NB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wi = wave & 1, wj = wave >> 1;  // wave's 32 x 32: rows i0+32wi.., cols j0+32wj..
  const int li = lane & 15, lk = lane >> 4;
  d4 sr[2][2], si[2][2];  // [j tile][i tile]
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) { sr[x][y] = d4{0, 0, 0, 0}; si[x][y] = d4{0, 0, 0, 0}; }
  // the tile's old values, loaded before the product so that their HBM
  // latency overlaps it (only the entries outside the pivot rows and columns
  // are read back)
  double2 old[2][2][4];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = i0 + 32 * wi + 16 * y + li, j = j0 + 32 * wj + 16 * x + 4 * v + lk;
        const bool keep = i < n && j < n && !(i >= p0 && i < p0 + nb) && !(j >= p0 && j < p0 + nb);
        old[x][y][v] = keep ? a[i + (size_t)n * j] : double2{0.0, 0.0};
      }
  for (int k0 = 0; k0 < nb; k0 += KH) {
    if (k0) __syncthreads();  // the previous half is consumed
    for (int e = tid; e < 64 * KH; e += 256) {
      const int r = e & 63, c = e >> 6;  // consecutive threads: consecutive rows (column-major L)
      const int i = i0 + r;
      const double2 v = (i < n && k0 + c < nb) ? l[i + (size_t)n * (k0 + c)] : double2{0.0, 0.0};
      Lr[c][r] = v.x;
      Li[c][r] = v.y;
    }
    for (int e = tid; e < KH * 64; e += 256) {
      const int r = e % KH, c = e / KH;  // consecutive threads: consecutive k of one column
      const int j = j0 + c, kr = k0 + r;
      double2 v = {0.0, 0.0};
      if (j < n && kr < nb) v = (j >= p0 && j < p0 + nb) ? b[kr + NB * (j - p0)] : u[kr + (size_t)NB * j];
      Ur[r][c] = v.x;
      Ui[r][c] = v.y;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < KH / 4; ks++) {
      const int k = 4 * ks + lk;
      double ur[2], ui[2], lr[2], lii[2], nli[2];
#pragma unroll
      for (int t = 0; t < 2; t++) {
        ur[t] = Ur[k][32 * wj + 16 * t + li];
        ui[t] = Ui[k][32 * wj + 16 * t + li];
        lr[t] = Lr[k][32 * wi + 16 * t + li];
        lii[t] = Li[k][32 * wi + 16 * t + li];
        nli[t] = -lii[t];
      }
#pragma unroll
      for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) {
          sr[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ur[x], lr[y], sr[x][y], 0, 0, 0);
          sr[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ui[x], nli[y], sr[x][y], 0, 0, 0);
          si[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ui[x], lr[y], si[x][y], 0, 0, 0);
          si[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ur[x], lii[y], si[x][y], 0, 0, 0);
        }
    }
  }
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int jl = 32 * wj + 16 * x + 4 * v + lk;  // D row
        const int il = 32 * wi + 16 * y + li;          // D column
        const int i = i0 + il, j = j0 + jl;
        if (i >= n || j >= n) continue;
        const bool ri = (i >= p0 && i < p0 + nb), cj = (j >= p0 && j < p0 + nb);
        double2* aij = a + i + (size_t)n * j;
        const double xr = sr[x][y][v], xi = si[x][y][v];
        if (ri) *aij = cj ? b[(i - p0) + NB * (j - p0)] : u[(i - p0) + (size_t)NB * j];  // B or U_j
        else if (cj) *aij = double2{-xr, -xi};                     // -L_i B
        else { const double2 o = old[x][y][v]; *aij = double2{o.x - xr, o.y - xi}; }  // A - L U
      }
}

// the strictly lower 64 x 64 tiles of the inverse, A[i, j] = conj(A[j, i]):
// one block per 32 x 32 tile of the strictly lower tile rows, staged through
// LDS so that reads and writes walk consecutive rows
__global__ void __launch_bounds__(256) pygw_ei_mirror(double2* __restrict__ A, int n) {
  __shared__ double2 X[32][33];
  const size_t w = blockIdx.z;
  double2* a = A + w * (size_t)n * n;
  const int i0 = 32 * blockIdx.x, j0 = 32 * blockIdx.y;  // target tile rows i0.., columns j0..
  if (i0 / 64 <= j0 / 64) return;                        // not strictly below the 64-tile diagonal
  const int tr = threadIdx.x & 31, tc = threadIdx.x >> 5;
  for (int c = tc; c < 32; c += 8) {
    const int i = j0 + tr, j = i0 + c;  // source A[j0 + tr][i0 + c]
    X[c][tr] = (i < n && j < n) ? a[i + (size_t)n * j] : double2{0.0, 0.0};
  }
  __syncthreads();
  for (int c = tc; c < 32; c += 8) {
    const int i = i0 + tr, j = j0 + c;  // A[i][j] = conj(A[j][i]) = conj(X[tr][c])
    const double2 v = X[tr][c];
    if (i < n && j < n) a[i + (size_t)n * j] = double2{v.x, -v.y};
  }
}

// (eps + eps^H)/2 + 1, inverted in place (the -1 is left to the caller)
int eps_inverse_gj(void* eps, int n, int nom, void* ws, int* dinfo, hipStream_t stream) {
  using namespace ei;
  double2* A = (double2*)eps;
  double2* B = (double2*)ws;
  double2* U = B + (size_t)nom * NB * NB;
  double2* L = U + (size_t)nom * NB * n;
  hipLaunchKernelGGL(pygw_ei_hermitize, dim3(ceil_div(n, 32), ceil_div(n, 32), nom), dim3(256), 0, stream, A, n);
  const int T = ceil_div(n, 64), ntu = T * (T + 1) / 2;   // 64 x 64 tiles on and above the diagonal
  const long per = ((long)nom * ntu + 7) / 8;
  for (int p0 = 0; p0 < n; p0 += NB) {
    const int nb = (n - p0) < NB ? (n - p0) : NB;
    hipLaunchKernelGGL(pygw_ei_pivot, dim3(nom), dim3(1024), 0, stream, A, n, p0, nb, B, dinfo);
    hipLaunchKernelGGL(pygw_ei_panels, dim3(ceil_div(n, 64), nom), dim3(256), 0, stream, A, n, p0, nb, B, U, L);
    hipLaunchKernelGGL(pygw_ei_update_mfma, dim3((unsigned)(8 * per)), dim3(256), 0, stream, A, n, p0, nb, B, U, L,
                       nom, ntu, per);
  }
  hipLaunchKernelGGL(pygw_ei_mirror, dim3(ceil_div(n, 32), ceil_div(n, 32), nom), dim3(256), 0, stream, A, n);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

int eps_minus_identity(void* eps, int n, int nom, hipStream_t stream) {
  hipLaunchKernelGGL(pygw_ei_minus_identity, dim3(ceil_div((long)n * nom, 256)), dim3(256), 0, stream, (double2*)eps, n,
                     nom);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

int eps_inverse_launch(void* eps, int n, int nom, void* ws, int* dinfo, hipStream_t stream) {
  const int rc = eps_inverse_gj(eps, n, nom, ws, dinfo, stream);
  return rc == PYGW_OK ? eps_minus_identity(eps, n, nom, stream) : rc;
}

size_t eps_inverse_ws_bytes(int n, int nom) {
  using namespace ei;
  return sizeof(double2) * ((size_t)nom * NB * NB + 2 * (size_t)nom * NB * n) + sizeof(int) * (size_t)nom + 256;
}
