// The orthogonalised plane-wave basis of the interstitial region and the
// plane-wave block of the Coulomb matrix, on gfx950:
//
//   OrthogonalizedBasisInterstitials (pygw.py:2126-2214, Real_sqrt_olap =
//   True, the cmp_pw_olap2 path, for_Coulomb.f90:306-325):
//     olap[j][i] = ipwint[G_i - G_j]  (i, j < ngq; the interstitial overlap)
//     tmat[j][i] = ipwint[G_i - G_j]  (j < ngq, i < ngq_barc)
//     mpwipw = olap^(-1/2) tmat, olap^(-1/2) = V diag(|e|^(-1/2)) V^H
//   with G_i = gindex[indgq[i]] and ipwint looked up through i_g0;
//
//   the plane-wave block of Coulomb (pygw.py:2419-2430):
//     Vmat[row0 + i][row0 + k] = sum_j mpwipw[i][j] 4 pi / gqlen[j]^2 conj(mpwipw[k][j])
//   (gqlen[0] taken as 1e100 at q = 0).
//
// The eigendecomposition of olap is the block Jacobi method of the Coulomb
// basis; the products run on the shared FP64 MFMA tile.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "gemm_tile.h"

size_t coul_setev_ws_bytes(int mbsize);
int coul_setev_launch(const void* vmat, int n0, double evtol, const void* wi0, int iop_coul_x, double* ev_out,
                      double* barcev, void* barcvm, void* vecs, int* matsize, int* immax, int* sweeps, void* ws,
                      size_t ws_size, hipStream_t stream);

// out[j][i] = ipwint[i_g0(G_i - G_j)] for j < nrow, i < ncol; 0 where the
// difference leaves the i_g0 box or has no index (cmp_pw_olap2 instead
// reuses the index of the previous pair there, for_Coulomb.f90:321-322)
__global__ void __launch_bounds__(256) pygw_pw_olap(const int* __restrict__ gindex, int ng,
                                                   const int* __restrict__ indgq, int nrow, int ncol,
                                                   const double2* __restrict__ ipwint,
                                                   const int* __restrict__ i_g0, int n1, int n2, int n3,
                                                   double2* __restrict__ out) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)nrow * ncol) return;
  const int j = (int)(e / ncol), i = (int)(e % ncol);
  const int a = indgq[i], b = indgq[j];
  double2 v{0.0, 0.0};
  if (a >= 0 && a < ng && b >= 0 && b < ng) {
    const int x = gindex[3 * a] - gindex[3 * b], y = gindex[3 * a + 1] - gindex[3 * b + 1],
              z = gindex[3 * a + 2] - gindex[3 * b + 2];
    if (abs(x) <= n1 && abs(y) <= n2 && abs(z) <= n3) {
      const int idg = i_g0[((size_t)(x + n1) * (2 * n2 + 1) + (y + n2)) * (2 * n3 + 1) + (z + n3)];
      if (idg >= 0 && idg < ng) v = ipwint[idg];
    }
  }
  out[e] = v;
}

// w[m] = |e_m|^(-3/2): olap^(-1/2) = sum_m b_m w_m b_m^H with b_m = v_m |e_m|^(1/2)
__global__ void pygw_pw_weights(const double* __restrict__ e, int n, double* __restrict__ w) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < n) {
    const double a = fabs(e[m]);
    w[m] = 1.0 / (a * sqrt(a));
  }
}

namespace pw {
template <class ST>
__device__ __forceinline__ void store(const d4 (&aR)[2][2], const d4 (&aI)[2][2], int m0, int n0, ST st) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave / is::L<2>::WC, wc = wave % is::L<2>::WC, li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int v = 0; v < 4; v++)
        st(m0 + 32 * wr + 16 * a + 4 * v + lk, n0 + 32 * wc + 16 * b + li, double2{aR[a][b][v], aI[a][b][v]});
}
}  // namespace pw

// S[r][k] = sum_m B[r + m n] w[m] conj(B[k + m n]) (B Fortran [n][n])
__global__ void __launch_bounds__(256) pygw_pw_isqrt(const double2* __restrict__ B, int n,
                                                    const double* __restrict__ w, double2* __restrict__ S) {
  using namespace is;
  __shared__ double2 As[64][KC + 1];
  __shared__ double2 Bs[KC][64 + 1];
  const int nt = (n + 63) / 64, m0 = (blockIdx.x / nt) * 64, n0 = (blockIdx.x % nt) * 64;
  d4 aR[2][2], aI[2][2];
  zero(aR, aI);
  auto la = [&](int m, int k) -> double2 {
    const int r = m0 + m;
    if (r >= n || k >= n) return double2{0.0, 0.0};
    const double2 x = B[r + (size_t)k * n];
    return double2{x.x * w[k], x.y * w[k]};
  };
  auto lb = [&](int k, int c) -> double2 {
    const int col = n0 + c;
    if (col >= n || k >= n) return double2{0.0, 0.0};
    const double2 x = B[col + (size_t)k * n];
    return double2{x.x, -x.y};
  };
  tile<2>(n, la, lb, As, Bs, aR, aI);
  pw::store(aR, aI, m0, n0, [&](int r, int c, double2 v) {
    if (r < n && c < n) S[(size_t)r * n + c] = v;
  });
}

// out[r][c] = sum_i S[r][i] T[i][c] (S [n][n], T and out [n][ncol])
__global__ void __launch_bounds__(256) pygw_pw_apply(const double2* __restrict__ S, int n,
                                                    const double2* __restrict__ T, int ncol,
                                                    double2* __restrict__ out) {
  using namespace is;
  __shared__ double2 As[64][KC + 1];
  __shared__ double2 Bs[KC][64 + 1];
  const int ntc = (ncol + 63) / 64, m0 = (blockIdx.x / ntc) * 64, n0 = (blockIdx.x % ntc) * 64;
  d4 aR[2][2], aI[2][2];
  zero(aR, aI);
  auto la = [&](int m, int k) -> double2 {
    const int r = m0 + m;
    return (r < n && k < n) ? S[(size_t)r * n + k] : double2{0.0, 0.0};
  };
  auto lb = [&](int k, int c) -> double2 {
    const int col = n0 + c;
    return (k < n && col < ncol) ? T[(size_t)k * ncol + col] : double2{0.0, 0.0};
  };
  tile<2>(n, la, lb, As, Bs, aR, aI);
  pw::store(aR, aI, m0, n0, [&](int r, int c, double2 v) {
    if (r < n && c < ncol) out[(size_t)r * ncol + c] = v;
  });
}

// V[(row0 + i) ld + row0 + k] = sum_j mp[i][j] vq_j conj(mp[k][j]), vq_j = 4 pi / gqlen_j^2
__global__ void __launch_bounds__(256) pygw_pw_vblock(const double2* __restrict__ mp, int n, int ncol,
                                                     const double* __restrict__ gqlen, int q_is_zero,
                                                     double2* __restrict__ V, long ld, int row0) {
  using namespace is;
  __shared__ double2 As[64][KC + 1];
  __shared__ double2 Bs[KC][64 + 1];
  const int nt = (n + 63) / 64, m0 = (blockIdx.x / nt) * 64, n0 = (blockIdx.x % nt) * 64;
  const double fourpi = 4.0 * 3.14159265358979323846;
  d4 aR[2][2], aI[2][2];
  zero(aR, aI);
  auto la = [&](int m, int k) -> double2 {
    const int r = m0 + m;
    if (r >= n || k >= ncol) return double2{0.0, 0.0};
    const double g = (q_is_zero && k == 0) ? 1e100 : gqlen[k];
    const double vq = fourpi / (g * g);
    const double2 x = mp[(size_t)r * ncol + k];
    return double2{x.x * vq, x.y * vq};
  };
  auto lb = [&](int k, int c) -> double2 {
    const int col = n0 + c;
    if (col >= n || k >= ncol) return double2{0.0, 0.0};
    const double2 x = mp[(size_t)col * ncol + k];
    return double2{x.x, -x.y};
  };
  tile<2>(ncol, la, lb, As, Bs, aR, aI);
  pw::store(aR, aI, m0, n0, [&](int r, int c, double2 v) {
    if (r < n && c < n) V[(size_t)(row0 + r) * ld + row0 + c] = v;
  });
}

namespace pw {
inline size_t al(size_t x) { return (x + 255) & ~(size_t)255; }
}  // namespace pw

size_t pw_basis_ws_bytes(int ngq, int ngq_barc) {
  using pw::al;
  const size_t n = ngq;
  return al(n * n * 16) * 3 + al(n * ngq_barc * 16) + al(n * 8) * 3 + al(coul_setev_ws_bytes(ngq));
}

int pw_basis_launch(const int* indgq, int ngq, int ngq_barc, const void* ipwint, int ng, const int* gindex,
                    const int* i_g0, int n1, int n2, int n3, void* mpwipw, void* ws, size_t ws_size,
                    hipStream_t stream) {
  using pw::al;
  if (ws_size < pw_basis_ws_bytes(ngq, ngq_barc)) return PYGW_E_ARG;
  const size_t n = ngq;
  char* p = (char*)ws;
  double2* olap = (double2*)p;
  p += al(n * n * 16);
  double2* B = (double2*)p;
  p += al(n * n * 16);
  double2* S = (double2*)p;
  p += al(n * n * 16);
  double2* T = (double2*)p;
  p += al(n * ngq_barc * 16);
  double* ev = (double*)p;
  p += al(n * 8);
  double* be = (double*)p;
  p += al(n * 8);
  double* w = (double*)p;
  p += al(n * 8);
  void* jws = p;
  hipLaunchKernelGGL(pygw_pw_olap, dim3(ceil_div((long)n * n, 256)), dim3(256), 0, stream, gindex, ng, indgq, ngq,
                     ngq, (const double2*)ipwint, i_g0, n1, n2, n3, olap);
  hipLaunchKernelGGL(pygw_pw_olap, dim3(ceil_div((long)n * ngq_barc, 256)), dim3(256), 0, stream, gindex, ng, indgq,
                     ngq, ngq_barc, (const double2*)ipwint, i_g0, n1, n2, n3, T);
  HIP_TRY(hipGetLastError());
  int ms = 0;
  // every eigenvector kept (evtol below any eigenvalue): B = V diag(|e|^(1/2))
  const int rc = coul_setev_launch(olap, ngq, -DBL_MAX, nullptr, 0, ev, be, B, nullptr, &ms, nullptr, nullptr, jws,
                                   coul_setev_ws_bytes(ngq), stream);
  if (rc != PYGW_OK) return rc;
  if (ms != ngq) return PYGW_E_NOTPD;
  // olap is a Gram matrix: a (numerically) zero eigenvalue means linearly
  // dependent plane waves (a repeated G in indgq), and olap^(-1/2) does not exist
  std::vector<double> evh(n);
  HIP_TRY(hipMemcpyAsync(evh.data(), ev, n * sizeof(double), hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  double emax = 0.0, emin = DBL_MAX;
  for (double e : evh) {
    emax = std::max(emax, std::fabs(e));
    emin = std::min(emin, e);
  }
  if (!(emin > (double)n * DBL_EPSILON * emax)) return PYGW_E_NOTPD;
  hipLaunchKernelGGL(pygw_pw_weights, dim3(ceil_div(ngq, 256)), dim3(256), 0, stream, be, ngq, w);
  const int nt = ceil_div(ngq, 64);
  hipLaunchKernelGGL(pygw_pw_isqrt, dim3(nt * nt), dim3(256), 0, stream, B, ngq, w, S);
  hipLaunchKernelGGL(pygw_pw_apply, dim3(nt * ceil_div(ngq_barc, 64)), dim3(256), 0, stream, S, ngq, T, ngq_barc,
                     (double2*)mpwipw);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

int coul_pw_block_launch(const void* mpwipw, int ngq, int ngq_barc, const double* gqlen, int q_is_zero, void* vmat,
                         long ld, int row0, hipStream_t stream) {
  const int nt = ceil_div(ngq, 64);
  hipLaunchKernelGGL(pygw_pw_vblock, dim3(nt * nt), dim3(256), 0, stream, (const double2*)mpwipw, ngq, ngq_barc,
                     gqlen, q_is_zero, (double2*)vmat, ld, row0);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}
