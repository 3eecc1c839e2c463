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
#include "common.h"

namespace ei {
constexpr int NB = 32;   // pivot block
constexpr int TT = 32;   // update tile edge
}

// (A + A^H)/2 + I, in place; one thread per (i <= j).
__global__ void __launch_bounds__(256) pygw_ei_hermitize(double2* __restrict__ A, int n) {
  const size_t w = blockIdx.y;
  double2* a = A + w * (size_t)n * n;
  const int j = blockIdx.x * 16 + threadIdx.y;
  for (int i = threadIdx.x; i <= j && j < n; i += 16) {
    double2 x = a[i + (size_t)n * j];
    double2 y = a[j + (size_t)n * i];
    double2 h = {(x.x + y.x) / 2.0, (x.y - y.y) / 2.0};
    if (i == j) { h.x = h.x + 1.0; h.y = 0.0; }
    a[i + (size_t)n * j] = h;
    a[j + (size_t)n * i] = double2{h.x, -h.y};
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
// One block of 256 threads per frequency.  info[w] |= 1 if a pivot is not
// positive (matrix not positive definite).
__global__ void __launch_bounds__(256) pygw_ei_pivot(const double2* __restrict__ A, int n, int p0, int nb,
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
  for (int k = 0; k < nb; k++) {
    const double2 piv = S[k][k];
    if (threadIdx.x == 0 && !(piv.x > 0.0)) atomicOr(&info[w], 1);
    __syncthreads();
    // update all (i, j) with i, j != k first, using the pre-step row/col k
    for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
      const int i = e % nb, j = e / nb;
      if (i != k && j != k) {
        const double2 t = cdiv(cmul(S[i][k], S[k][j]), piv);
        S[i][j].x -= t.x;
        S[i][j].y -= t.y;
      }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nb; e += blockDim.x) {
      if (e != k) {
        const double2 c = cdiv(S[e][k], piv);
        const double2 r = cdiv(S[k][e], piv);
        S[e][k] = double2{-c.x, -c.y};
        S[k][e] = r;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) S[k][k] = cdiv(double2{1.0, 0.0}, piv);
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
// grid (ceil(n/64), nom), 64 threads: thread = one column j (for U) and one
// row i (for L).
__global__ void __launch_bounds__(64) pygw_ei_panels(const double2* __restrict__ A, int n, int p0, int nb,
                                                     const double2* __restrict__ B, double2* __restrict__ U,
                                                     double2* __restrict__ L) {
  using namespace ei;
  __shared__ double2 Bs[NB][NB];
  const size_t w = blockIdx.y;
  const double2* a = A + w * (size_t)n * n;
  const double2* b = B + w * NB * NB;
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) Bs[e % nb][e / nb] = b[(e % nb) + NB * (e / nb)];
  __syncthreads();
  const int j = blockIdx.x * 64 + threadIdx.x;
  if (j >= n) return;
  double2 col[NB];
  for (int r = 0; r < nb; r++) col[r] = a[(p0 + r) + (size_t)n * j];
  double2* u = U + w * (size_t)NB * n;
  for (int r = 0; r < nb; r++) {
    double2 s = {0.0, 0.0};
    for (int c = 0; c < nb; c++) {
      const double2 t = cmul(Bs[r][c], col[c]);
      s.x += t.x;
      s.y += t.y;
    }
    u[r + (size_t)NB * j] = s;
  }
  double2* l = L + w * (size_t)n * NB;
  for (int c = 0; c < nb; c++) l[j + (size_t)n * c] = a[j + (size_t)n * (p0 + c)];
}

// Step 3: rank-nb update of all TT x TT tiles.
//   rows i in P, cols j in P : A_PP = B
//   rows i in P              : A_Pj = U_j
//   cols j in P              : A_iP = -L_i B
//   otherwise                : A_ij -= L_i U_j
__global__ void __launch_bounds__(256) pygw_ei_update(double2* __restrict__ A, int n, int p0, int nb,
                                                      const double2* __restrict__ B, const double2* __restrict__ U,
                                                      const double2* __restrict__ L) {
  using namespace ei;
  __shared__ double2 Ls[TT][NB + 1];
  __shared__ double2 Us[NB][TT + 1];
  const size_t w = blockIdx.z;
  double2* a = A + w * (size_t)n * n;
  const double2* u = U + w * (size_t)NB * n;
  const double2* l = L + w * (size_t)n * NB;
  const double2* b = B + w * NB * NB;
  const int i0 = blockIdx.x * TT, j0 = blockIdx.y * TT;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  // stage L rows i0.., U cols j0.. (for P-columns, U_j is replaced by B)
  for (int e = threadIdx.x; e < TT * NB; e += 256) {
    const int r = e % TT, c = e / TT;
    const int i = i0 + r;
    Ls[r][c] = (i < n && c < nb) ? l[i + (size_t)n * c] : double2{0.0, 0.0};
  }
  for (int e = threadIdx.x; e < NB * TT; e += 256) {
    const int r = e % NB, c = e / NB;
    const int j = j0 + c;
    double2 v = {0.0, 0.0};
    if (j < n && r < nb) v = (j >= p0 && j < p0 + nb) ? b[r + NB * (j - p0)] : u[r + (size_t)NB * j];
    Us[r][c] = v;
  }
  __syncthreads();
  for (int q = 0; q < TT / 8; q++) {
    const int i = i0 + tx, j = j0 + ty + 8 * q;
    if (i >= n || j >= n) continue;
    const bool ri = (i >= p0 && i < p0 + nb), cj = (j >= p0 && j < p0 + nb);
    double2 s = {0.0, 0.0};
    for (int c = 0; c < nb; c++) {
      const double2 t = cmul(Ls[tx][c], Us[c][ty + 8 * q]);
      s.x += t.x;
      s.y += t.y;
    }
    double2* aij = a + i + (size_t)n * j;
    if (ri && cj) *aij = Us[i - p0][ty + 8 * q];                     // B
    else if (ri) *aij = Us[i - p0][ty + 8 * q];                      // U_j
    else if (cj) *aij = double2{-s.x, -s.y};                         // -L_i B
    else { double2 o = *aij; *aij = double2{o.x - s.x, o.y - s.y}; } // A - L U
  }
}

int eps_inverse_launch(void* eps, int n, int nom, void* ws, int* dinfo, hipStream_t stream) {
  using namespace ei;
  double2* A = (double2*)eps;
  double2* B = (double2*)ws;
  double2* U = B + (size_t)nom * NB * NB;
  double2* L = U + (size_t)nom * NB * n;
  hipLaunchKernelGGL(pygw_ei_hermitize, dim3(ceil_div(n, 16), nom), dim3(16, 16), 0, stream, A, n);
  for (int p0 = 0; p0 < n; p0 += NB) {
    const int nb = (n - p0) < NB ? (n - p0) : NB;
    hipLaunchKernelGGL(pygw_ei_pivot, dim3(nom), dim3(256), 0, stream, A, n, p0, nb, B, dinfo);
    hipLaunchKernelGGL(pygw_ei_panels, dim3(ceil_div(n, 64), nom), dim3(64), 0, stream, A, n, p0, nb, B, U, L);
    hipLaunchKernelGGL(pygw_ei_update, dim3(ceil_div(n, TT), ceil_div(n, TT), nom), dim3(256), 0, stream, A, n, p0,
                       nb, B, U, L);
  }
  hipLaunchKernelGGL(pygw_ei_minus_identity, dim3(ceil_div((long)n * nom, 256)), dim3(256), 0, stream, A, n, nom);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

size_t eps_inverse_ws_bytes(int n, int nom) {
  using namespace ei;
  return sizeof(double2) * ((size_t)nom * NB * NB + 2 * (size_t)nom * NB * n) + sizeof(int) * (size_t)nom + 256;
}
