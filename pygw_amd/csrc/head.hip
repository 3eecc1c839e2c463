// The head of the dielectric matrix at q = 0 and the momentum matrix elements
// it is built from, on gfx950 (calc_head, pygw.py:2808-3037).
//
//   mmatvv[ie2 - mst][ie1 - nst][c] = (p + conj(p'))/2, p = <ie1| -i nabla |ie2>,
//   p' = <ie2| -i nabla |ie1>, each the muffin-tin part (calcmmatvv_MT,
//   for_q_0.f90:10-234) plus the interstitial part (calcmmatvv_is,
//   for_q_0.f90:460-507).  Muffin tin: the coupling of (l, m) to (l +- 1,
//   m + {-1, 0, 1}) applied to every band's coefficients (one thread per
//   (band, channel)), then sum over channels of conj(u_n1) (K u_n2) in both
//   orders as FP64-MFMA products.  Interstitial: the overlap O[G2][G1] =
//   ipwint(G2 - G1) gathered through i_g0, T[G2][b] = sum_G1 O conj(A[b][G1])
//   as an FP64-MFMA product, then one wave per pair for sum_G2 A[j][G2]
//   (k+G2)_c T[G2][i] in both orders.  The products split their sums over
//   blocks (small band windows leave few output tiles) and add the chunks in
//   a fixed order.
//   mmatcv[ie2 - mst][icg][c] (calcmmatcv, for_q_0.f90:510-683): one thread
//   per (band, core state).
//   head[iom] = -sum_irk 4 pi fspin w_irk/Vol sum |p|^2/3/(E_c - E_v)^2 kcw
//   (pygw.py:2940-3009): the terms once per (irk, pair), then one block per
//   frequency contracting them with kcw's contiguous (pair) slab of each k;
//   the metallic intraband sum and the Drude term (pygw.py:2952-2959,
//   3015-3025) in a single-block pass.
//   pm (pygw.py:3122-3141): one thread per (k, pair).
#include "gemm_tile.h"

#include <algorithm>
#include <cmath>

namespace hd {
constexpr int kMaxNat = 16;
constexpr int kMaxRows = 17;  // rows of nLO_at (lomax + 1)

constexpr int kMaxNdf = 64;

struct Layout {
  int nat, ndf, lmax, ntnt, nbmax, nbmax_g, nlomax, nlo, lox, nrows;
  int mult[kMaxNat];
  int nlo_at[kMaxNat * kMaxRows];  // nLO_at[l + nrows iat]
  int core_off[kMaxNat];
  int df_atom[kMaxNdf];            // atom type of each atom
};

struct VvTables {  // [2][lmax][nat] and [2][nlo][lox][nat], [2][nlo][nlo][lox][nat] (Fortran)
  const double *ul_ul, *ul_udl, *udl_ul, *udl_udl, *ulol_ul, *ulol_udl, *ul_ulol, *udl_ulol, *ulol_ulol;
};

struct CvTables {  // [2][nct] and [2][nloat][nct] (Fortran)
  const double *ul_ucl, *udl_ucl, *ucl_ul, *ucl_udl, *ulol_ucl, *ucl_ulol;
  int nloat, nct;
};

__device__ inline double2 cmul(double2 a, double2 b) { return double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ inline double2 cadd(double2 a, double2 b) { return double2{a.x + b.x, a.y + b.y}; }
__device__ inline double2 rmul(double2 a, double r) { return double2{a.x * r, a.y * r}; }
__device__ inline double2 conjd(double2 a) { return double2{a.x, -a.y}; }

__device__ inline int nlo_of(const Layout& L, int l, int iat) {
  return (l >= 0 && l < L.nrows) ? L.nlo_at[l + L.nrows * iat] : 0;
}

struct Bands {  // Fortran alfa/beta (nbmax, ntnt, ndf), gama (nbmax_g, nlomax, ntnt, ndf)
  const double2 *a, *b, *g;
  int nbmax, nbmax_g, ntnt, nlomax;
  __device__ double2 A(int n, int t, int d) const { return a[n + (size_t)nbmax * (t + (size_t)ntnt * d)]; }
  __device__ double2 B(int n, int t, int d) const { return b[n + (size_t)nbmax * (t + (size_t)ntnt * d)]; }
  __device__ double2 G(int n, int o, int t, int d) const {
    return g[n + (size_t)nbmax_g * (o + (size_t)nlomax * (t + (size_t)ntnt * d))];
  }
};

}  // namespace hd

// The spheres' part as a sparse operator and two products.  With u_n[ch]
// the coefficients of band n over the channels ch = (atom, (l, m), kind:
// u, udot or local orbital o < nLO_at(l)), p_c(n1, n2) = sum_ch conj(u_n1[ch])
// V_c,n2[ch], where V_c,n2 = K_c u_n2 collects, for the bra channel ch =
// (L, M, X), the ket channels (L -/+ 1, M + {-1, 0, 1}, Y) of the six sums
// of for_q_0.f90:236-458 weighted by their Clebsch-Gordan factor and radial
// integral (R1 = <X_l+1| d/dr - l/r |Y_l>, R2 = <X_l| d/dr + (l+2)/r |Y_l+1>).
// Thread per (band, channel): Uc[n][ch] = conj(u_n[ch]) and V[c][n][ch].
namespace hd {
struct ChanMap {  // channel ch = t + ntnt (kind + nk idf), kind 0 u, 1 udot, 2 + o local orbital o
  int ntnt, nk;
};

// radial integral of the momentum operator between bra kind kx (at lb) and
// ket kind ky: il = 0 the R1 table at l (bra at l+1), il = 1 the R2 table
__device__ inline double rad(const Layout& L, const VvTables& T, int il, int l, int iat, int kx, int ky) {
  auto t2 = [&](const double* t) { return t[il + 2 * (l + L.lmax * iat)]; };
  auto t4 = [&](const double* t, int o) { return t[il + 2 * (o + L.nlo * (l + L.lox * iat))]; };
  if (kx < 2 && ky < 2) return kx == 0 ? (ky == 0 ? t2(T.ul_ul) : t2(T.ul_udl)) : (ky == 0 ? t2(T.udl_ul) : t2(T.udl_udl));
  if (kx < 2) return t4(kx == 0 ? T.ul_ulol : T.udl_ulol, ky - 2);
  if (ky < 2) return t4(ky == 0 ? T.ulol_ul : T.ulol_udl, kx - 2);
  return T.ulol_ulol[il + 2 * ((kx - 2) + L.nlo * ((ky - 2) + L.nlo * (l + L.lox * iat)))];
}

__device__ inline double2 coef(const Bands& C, int n, int t, int kind, int idf) {
  return kind == 0 ? C.A(n, t, idf) : (kind == 1 ? C.B(n, t, idf) : C.G(n, kind - 2, t, idf));
}
}  // namespace hd

__global__ void __launch_bounds__(256) pygw_hd_kapply(hd::Layout L, hd::VvTables T, hd::Bands C, int b0, int nbw,
                                                     long nch, double2* __restrict__ Uc, double2* __restrict__ V) {
  using namespace hd;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= nbw * nch) return;
  const int nb = (int)(e / nch), n = b0 + nb;
  const long ch = e % nch;
  const int nk = 2 + L.nlomax;
  const int t = (int)(ch % L.ntnt), kind = (int)((ch / L.ntnt) % nk), idf = (int)(ch / ((long)L.ntnt * nk));
  const int iat = L.df_atom[idf];
  const int Lb = (int)sqrtf((float)t + 0.5f), Mb = t - Lb * (Lb + 1);
  const bool valid = kind < 2 || (kind - 2) < nlo_of(L, Lb, iat);
  const double2 z{0.0, 0.0};
  Uc[e] = valid ? conjd(coef(C, n, t, kind, idf)) : z;
  double2 pxpy = z, pxmy = z, pz = z;
  if (valid) {
    // ket sum over kinds Y at (l, m) with radial table (il, lr)
    auto ket = [&](int il, int lr, int l, int m) {
      double2 s = z;
      const int tk = l * (l + 1) + m, nl = nlo_of(L, l, iat);
      for (int ky = 0; ky < 2 + nl; ky++) s = cadd(s, rmul(coef(C, n, tk, ky, idf), rad(L, T, il, lr, iat, kind, ky)));
      return s;
    };
    if (Lb >= 1 && Lb - 1 < L.lmax) {  // bra at l + 1 = Lb: xpy1 (m = Mb - 1), xmy1 (m = Mb + 1), z1 (m = Mb)
      const int l = Lb - 1;
      const double den = (double)((2 * l + 1) * (2 * l + 3));
      int m = Mb - 1;
      if (m >= -l && m <= l) pxpy = cadd(pxpy, rmul(ket(0, l, l, m), -sqrt((double)((l + m + 1) * (l + m + 2)) / den)));
      m = Mb + 1;
      if (m >= -l && m <= l) pxmy = cadd(pxmy, rmul(ket(0, l, l, m), sqrt((double)((l - m + 1) * (l - m + 2)) / den)));
      m = Mb;
      if (m >= -l && m <= l) pz = cadd(pz, rmul(ket(0, l, l, m), sqrt((double)((l - m + 1) * (l + m + 1)) / den)));
    }
    if (Lb < L.lmax) {  // bra at l = Lb: xpy2 (ket m - 1), xmy2 (ket m + 1), z2 (ket m) at l + 1
      const int l = Lb, m = Mb;
      const double den = (double)((2 * l + 1) * (2 * l + 3));
      pxpy = cadd(pxpy, rmul(ket(1, l, l + 1, m - 1), sqrt((double)((l - m + 1) * (l - m + 2)) / den)));
      pxmy = cadd(pxmy, rmul(ket(1, l, l + 1, m + 1), -sqrt((double)((l + m + 1) * (l + m + 2)) / den)));
      pz = cadd(pz, rmul(ket(1, l, l + 1, m), sqrt((double)((l - m + 1) * (l + m + 1)) / den)));
    }
  }
  // (-(pxpy + pxmy) i/2, (pxmy - pxpy)/2, -i pz)
  const double2 sm = cadd(pxpy, pxmy), df = cadd(pxmy, rmul(pxpy, -1.0));
  const size_t plane = (size_t)nbw * nch;
  V[e] = double2{0.5 * sm.y, -0.5 * sm.x};
  V[plane + e] = rmul(df, 0.5);
  V[2 * plane + e] = double2{pz.y, -pz.x};
}

// Complex products with both operands contiguous along the sum, split over
// the sum: part[s][job][m][n] = sum_{k in chunk s} A_job[m][k] B_job[n][k]
// (B conjugated with conjB).  Tiles of 32 x 32 on v_mfma_f64_16x16x4f64,
// both operands staged along k; the chunks are summed in a fixed order
// afterwards, so the result does not depend on the schedule.
namespace hd {
constexpr int kMaxJobs = 6;
struct Job {
  const double2 *A, *B;
  long lda, ldb, off;  // off: the job's first element in a partial plane
  int M, N, tiles_n, tile0;
};
struct Jobs {
  Job j[kMaxJobs];
  int njob, ntile, conjB;
  long K, chunk, plane;  // plane = sum over jobs of M N
};
}  // namespace hd

__global__ void __launch_bounds__(256) pygw_hd_gemm(hd::Jobs J, double2* __restrict__ part) {
  using namespace hd;
  constexpr int TT = 32, KC = 32;
  __shared__ double2 As[TT][KC + 1];
  __shared__ double2 Bs[TT][KC + 1];
  const int tile = blockIdx.x, sp = blockIdx.y;
  int jb = 0;
  while (jb + 1 < J.njob && tile >= J.j[jb + 1].tile0) jb++;
  const Job& Jb = J.j[jb];
  const int tl = tile - Jb.tile0;
  const int m0 = (tl / Jb.tiles_n) * TT, c0 = (tl % Jb.tiles_n) * TT;
  const long kb = sp * J.chunk, ke = std::min(J.K, kb + J.chunk);
  const double cs = J.conjB ? -1.0 : 1.0;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1, li = lane & 15, lk = lane >> 4;
  d4 aR = d4{0, 0, 0, 0}, aI = d4{0, 0, 0, 0};
  for (long k0 = kb; k0 < ke; k0 += KC) {
    for (int e = tid; e < TT * KC; e += 256) {
      const int c = e % KC, m = e / KC;
      const long k = k0 + c;
      double2 a{0.0, 0.0}, b{0.0, 0.0};
      if (k < ke) {
        if (m0 + m < Jb.M) a = Jb.A[(size_t)(m0 + m) * Jb.lda + k];
        if (c0 + m < Jb.N) b = Jb.B[(size_t)(c0 + m) * Jb.ldb + k];
      }
      As[m][c] = a;
      Bs[m][c] = double2{b.x, cs * b.y};
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < KC / 4; ks++) {
      const int kk = ks * 4 + lk;
      const double2 x = As[16 * wr + li][kk], y = Bs[16 * wc + li][kk];
      aR = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, y.x, aR, 0, 0, 0);
      aR = __builtin_amdgcn_mfma_f64_16x16x4f64(-x.y, y.y, aR, 0, 0, 0);
      aI = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, y.y, aI, 0, 0, 0);
      aI = __builtin_amdgcn_mfma_f64_16x16x4f64(x.y, y.x, aI, 0, 0, 0);
    }
    __syncthreads();
  }
  double2* o = part + (size_t)sp * J.plane + Jb.off;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int r = m0 + 16 * wr + 4 * v + lk, c = c0 + 16 * wc + li;
    if (r < Jb.M && c < Jb.N) o[(size_t)r * Jb.N + c] = double2{aR[v], aI[v]};
  }
}

// host: the jobs' tiles and the chunk of the sum that fills the chip
static hd::Jobs make_jobs(int njob, const hd::Job* jobs, long K, int conjB) {
  hd::Jobs J{};
  J.njob = njob;
  J.K = K;
  J.conjB = conjB;
  long off = 0;
  int tile = 0;
  for (int i = 0; i < njob; i++) {
    J.j[i] = jobs[i];
    J.j[i].off = off;
    J.j[i].tiles_n = ceil_div(jobs[i].N, 32);
    J.j[i].tile0 = tile;
    tile += ceil_div(jobs[i].M, 32) * J.j[i].tiles_n;
    off += (long)jobs[i].M * jobs[i].N;
  }
  J.ntile = tile;
  J.plane = off;
  // about 2 blocks per CU, chunks of whole KC steps, at most 32 of them
  int split = std::max(1, std::min(32, ceil_div(512, std::max(tile, 1))));
  split = std::min<long>(split, std::max(1L, K / 64));
  J.chunk = ((K + split - 1) / split + 31) / 32 * 32;
  return J;
}

static int n_split(const hd::Jobs& J) { return (int)((J.K + J.chunk - 1) / J.chunk); }

// mmatvv[ie2][ie1][c] = (P1_c[ie1][ie2] + conj(P2_c[ie2][ie1]))/2, the
// chunks of both products summed in order (jobs 0-2: P1_c [n1][n2], 3-5: P2_c
// [n2][n1])
__global__ void __launch_bounds__(256) pygw_hd_mtcombine(const double2* __restrict__ part, int nsplit, long plane,
                                                        int n1, int n2, double2* __restrict__ mmatvv) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n1 * n2 * 3) return;
  const int c = e % 3, ie1 = (e / 3) % n1, ie2 = e / (3 * n1);
  const size_t ip = (size_t)c * n1 * n2 + (size_t)ie1 * n2 + ie2;
  const size_t iq = (size_t)(3 + c) * n1 * n2 + (size_t)ie2 * n1 + ie1;
  double2 p{0.0, 0.0}, q{0.0, 0.0};
  for (int sp = 0; sp < nsplit; sp++) {
    const double2 x = part[sp * plane + ip], y = part[sp * plane + iq];
    p.x += x.x;
    p.y += x.y;
    q.x += y.x;
    q.y += y.y;
  }
  mmatvv[e] = double2{0.5 * (p.x + q.x), 0.5 * (p.y - q.y)};
}

// O[G2][G1] = ipwint(G2 - G1) over the k-point's plane waves (0 where the
// difference is outside i_g0 or i_g0 holds no row)
__global__ void __launch_bounds__(256) pygw_hd_olap(const int* __restrict__ indgk, int nvk,
                                                   const int* __restrict__ gindex, int ng,
                                                   const double2* __restrict__ ipwint, const int* __restrict__ i_g0,
                                                   int g1, int g2, int g3, double2* __restrict__ O) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)nvk * nvk) return;
  const int r = (int)(e / nvk), k = (int)(e % nvk);
  const int a = indgk[r], b = indgk[k];
  const int d0 = gindex[a] - gindex[b] + g1, d1 = gindex[a + ng] - gindex[b + ng] + g2,
            d2 = gindex[a + 2 * ng] - gindex[b + 2 * ng] + g3;
  double2 v{0.0, 0.0};
  if (d0 >= 0 && d0 <= 2 * g1 && d1 >= 0 && d1 <= 2 * g2 && d2 >= 0 && d2 <= 2 * g3) {
    const int idg = i_g0[d0 + (2 * g1 + 1) * (d1 + (2 * g2 + 1) * d2)];
    if (idg >= 0 && idg < ng) v = ipwint[idg];
  }
  O[e] = v;
}

// T[G2][b] = the chunks of O conj(A)^T summed in order
__global__ void __launch_bounds__(256) pygw_hd_sum(const double2* __restrict__ part, int nsplit, long plane,
                                                  double2* __restrict__ T) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= plane) return;
  double2 s{0.0, 0.0};
  for (int sp = 0; sp < nsplit; sp++) {
    const double2 x = part[sp * plane + e];
    s.x += x.x;
    s.y += x.y;
  }
  T[e] = s;
}

struct KpG {
  double kcar[3], k2c[9];
};

// one wave per pair: mmatvv += (p_is + conj(p'_is))/2,
// p_is(i, j) = sum_G2 A[j][G2] (k+G2) T[G2][i - b0]
__global__ void __launch_bounds__(256) pygw_hd_is_p(const double2* __restrict__ A, long lda, int b0, int nbw,
                                                   const double2* __restrict__ Tm, const int* __restrict__ indgk,
                                                   int nvk, const int* __restrict__ gindex, int ng, KpG kp, int nst,
                                                   int n1, int mst, int n2, double2* __restrict__ mmatvv) {
  const int lane = threadIdx.x & 63;
  const int pair = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (pair >= n1 * n2) return;
  const int ie1 = nst + pair % n1, ie2 = mst + pair / n1;
  double2 p[3] = {{0, 0}, {0, 0}, {0, 0}}, q[3] = {{0, 0}, {0, 0}, {0, 0}};
  for (int g = lane; g < nvk; g += 64) {
    const int a = indgk[g];
    const double G[3] = {(double)gindex[a], (double)gindex[a + ng], (double)gindex[a + 2 * ng]};
    double kg[3];
    for (int c = 0; c < 3; c++) kg[c] = kp.kcar[c] + (kp.k2c[c] * G[0] + kp.k2c[c + 3] * G[1] + kp.k2c[c + 6] * G[2]);
    const double2 t1 = Tm[(size_t)g * nbw + (ie1 - b0)], t2 = Tm[(size_t)g * nbw + (ie2 - b0)];
    const double2 x = hd::cmul(A[(size_t)ie2 * lda + g], t1), y = hd::cmul(A[(size_t)ie1 * lda + g], t2);
    for (int c = 0; c < 3; c++) {
      p[c] = hd::cadd(p[c], hd::rmul(x, kg[c]));
      q[c] = hd::cadd(q[c], hd::rmul(y, kg[c]));
    }
  }
  for (int off = 32; off > 0; off >>= 1)
    for (int c = 0; c < 3; c++) {
      p[c].x += __shfl_xor(p[c].x, off);
      p[c].y += __shfl_xor(p[c].y, off);
      q[c].x += __shfl_xor(q[c].x, off);
      q[c].y += __shfl_xor(q[c].y, off);
    }
  if (lane == 0) {
    double2* o = mmatvv + ((size_t)(ie2 - mst) * n1 + (ie1 - nst)) * 3;
    for (int c = 0; c < 3; c++) o[c] = double2{o[c].x + 0.5 * (p[c].x + q[c].x), o[c].y + 0.5 * (p[c].y - q[c].y)};
  }
}

// one thread per (band ie2, core state icg): mmatcv[ie2 - mst][icg][c]
// (calcmmatcv of the core state's atom; a row whose l + 1 exceeds lmax, or
// whose atom or core index is not in the tables, reads as NaN)
__global__ void __launch_bounds__(256) pygw_hd_cv(hd::Layout L, hd::CvTables T, hd::Bands C,
                                                 const int* __restrict__ corind, int ncg, int mst, int n2, int exact,
                                                 double2* __restrict__ mmatcv) {
  using namespace hd;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ncg * n2) return;
  const int icg = e % ncg, iv = mst + e / ncg;
  const int iat = corind[icg], idf = corind[icg + ncg], ic = corind[icg + 2 * ncg], l = corind[icg + 3 * ncg],
            m = corind[icg + 4 * ncg];
  double2* o = mmatcv + ((size_t)(e / ncg) * ncg + icg) * 3;
  int ndf = 0;
  for (int a = 0; a < L.nat; a++) ndf += L.mult[a];
  const int cidx = (iat >= 0 && iat < L.nat) ? L.core_off[iat] + ic : -1;
  if (iat < 0 || iat >= L.nat || idf < 0 || idf >= ndf || ic < 0 || cidx >= T.nct || l < 0 || l + 1 > L.lmax ||
      m < -l || m > l) {
    const double nan = __builtin_nan("");
    for (int c = 0; c < 3; c++) o[c] = double2{nan, nan};
    return;
  }
  auto t2 = [&](const double* t, int il) { return t[il + 2 * cidx]; };
  auto t3 = [&](const double* t, int il, int lo) { return t[il + 2 * (lo + T.nloat * cidx)]; };
  const double d1 = (double)((2 * l + 1) * (2 * l + 3)), d2 = (double)((2 * l - 1) * (2 * l + 1));
  const int lm1mm1 = (l - 1) * l + m - 1, lm1m = lm1mm1 + 1, lm1m1 = lm1mm1 + 2;
  const int l1mm1 = (l + 1) * (l + 2) + m - 1, l1m = l1mm1 + 1, l1m1 = l1mm1 + 2;
  const double flm1 = -sqrt((double)((l + m + 1) * (l + m + 2)) / d1);
  const double flm2 = sqrt((double)((l - m - 1) * (l - m)) / d2);
  const double flm3 = sqrt((double)((l - m + 1) * (l - m + 2)) / d1);
  const double flm4 = -sqrt((double)((l + m - 1) * (l + m)) / d2);
  const double flm5 = sqrt((double)((l - m + 1) * (l + m + 1)) / d1);
  const double flm6 = sqrt((double)((l - m) * (l + m)) / d2);
  const int nlm1 = l > 0 ? nlo_of(L, l - 1, iat) : 0, nlp1 = nlo_of(L, l + 1, iat);
  const double2 z{0, 0};
  // <core| p |psi_iv>
  double2 xpy1 = z, xmy1 = z, z1 = z;
  auto kv = [&](int t, int il) {  // A u + B udot (+ C ulo) of band iv against the core state
    double2 s = cadd(rmul(C.A(iv, t, idf), t2(T.ucl_ul, il)), rmul(C.B(iv, t, idf), t2(T.ucl_udl, il)));
    const int nl = il == 0 ? nlm1 : nlp1;
    for (int o2 = 0; o2 < nl; o2++) s = cadd(s, rmul(C.G(iv, o2, t, idf), t3(T.ucl_ulol, il, o2)));
    return s;
  };
  if (l > 0 && l + m - 2 >= 0) xpy1 = kv(lm1mm1, 0);
  if (l > 0 && l - m - 2 >= 0) xmy1 = kv(lm1m1, 0);
  if (l > 0 && l + m - 1 >= 0 && l - m - 1 >= 0) z1 = kv(lm1m, 0);
  const double2 xpy2 = kv(l1mm1, 1), xmy2 = kv(l1m1, 1), z2 = kv(l1m, 1);
  double2 pxpy = cadd(rmul(xpy1, flm4), rmul(xpy2, flm3));
  double2 pxmy = cadd(rmul(xmy1, flm2), rmul(xmy2, flm1));
  double2 pz = cadd(rmul(z1, flm6), rmul(z2, flm5));
  // mcv = (-(pxpy + pxmy) i/2, -(pxpy - pxmy)/2, -i pz)
  double2 s = cadd(pxpy, pxmy), d = cadd(pxpy, rmul(pxmy, -1.0));
  const double2 mx_cv{0.5 * s.y, -0.5 * s.x}, my_cv = rmul(d, -0.5), mz_cv{pz.y, -pz.x};
  // <psi_iv| p |core>
  auto bv = [&](int t, int il, bool typo) {
    double2 sum = cadd(rmul(conjd(C.A(iv, t, idf)), t2(T.ul_ucl, il)), rmul(conjd(C.B(iv, t, idf)), t2(T.udl_ucl, il)));
    const int nl = il == 0 ? nlp1 : nlm1;
    for (int o2 = 0; o2 < nl; o2++)
      sum = cadd(sum, rmul(conjd(C.G(iv, o2, t, idf)), typo ? t2(T.udl_ucl, il) : t3(T.ulol_ucl, il, o2)));
    return sum;
  };
  const double2 vxpy1 = bv(l1m1, 0, false), vxmy1 = bv(l1mm1, 0, false), vz1 = bv(l1m, 0, false);
  double2 vxpy2 = z, vxmy2 = z, vz2 = z;
  if (l > 0 && l - m - 2 >= 0) vxpy2 = bv(lm1m1, 1, false);
  // the reference multiplies the l - 1 local orbitals here by iudl_ucl (for_q_0.f90:648)
  if (l > 0 && l + m - 2 >= 0) vxmy2 = bv(lm1mm1, 1, !exact);
  if (l > 0 && l + m - 1 >= 0 && l - m - 1 >= 0) vz2 = bv(lm1m, 1, false);
  pxpy = cadd(rmul(vxpy1, flm1), rmul(vxpy2, flm2));
  pxmy = cadd(rmul(vxmy1, flm3), rmul(vxmy2, flm4));
  pz = cadd(rmul(vz1, flm5), rmul(vz2, flm6));
  s = cadd(pxpy, pxmy);
  d = cadd(pxpy, rmul(pxmy, -1.0));
  const double2 mx_vc{0.5 * s.y, -0.5 * s.x}, my_vc = rmul(d, -0.5), mz_vc{pz.y, -pz.x};
  o[0] = double2{0.5 * (mx_cv.x + mx_vc.x), 0.5 * (mx_cv.y - mx_vc.y)};
  o[1] = double2{0.5 * (my_cv.x + my_vc.x), 0.5 * (my_cv.y - my_vc.y)};
  o[2] = double2{0.5 * (mz_cv.x + mz_vc.x), 0.5 * (mz_cv.y - mz_vc.y)};
}

// term[irk][ie2 n_o + io] = |p|^2/3/(E_c - E)^2 (io < ncg_p: core state io;
// else valence band io - ncg_p, 0 where |E_c - E_v| < 1e-5)
__global__ void __launch_bounds__(256) pygw_hd_terms(const double2* __restrict__ mmatvv,
                                                    const double2* __restrict__ mmatcv, const double* __restrict__ Ebnd,
                                                    int ldE, const double* __restrict__ eig_core, int nirk, int nvbm,
                                                    int ncg, int ncg_p, int n_e, int numin, double* __restrict__ term) {
  const long n_o = ncg_p + nvbm;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= nirk * n_o * n_e) return;
  const int io = (int)(e % n_o), ie2 = (int)((e / n_o) % n_e), irk = (int)(e / (n_o * n_e));
  const double ec = Ebnd[(size_t)irk * ldE + numin + ie2];
  double p2, t = 0.0;
  if (io < ncg_p) {
    const double2* m = mmatcv + (((size_t)irk * n_e + ie2) * ncg + io) * 3;
    p2 = (m[0].x * m[0].x + m[0].y * m[0].y + m[1].x * m[1].x + m[1].y * m[1].y + m[2].x * m[2].x + m[2].y * m[2].y) / 3.0;
    const double d = ec - eig_core[io];
    t = p2 / (d * d);
  } else {
    const int ie1 = io - ncg_p;
    const double2* m = mmatvv + (((size_t)irk * n_e + ie2) * nvbm + ie1) * 3;
    p2 = (m[0].x * m[0].x + m[0].y * m[0].y + m[1].x * m[1].x + m[1].y * m[1].y + m[2].x * m[2].x + m[2].y * m[2].y) / 3.0;
    const double d = ec - Ebnd[(size_t)irk * ldE + ie1];
    t = fabs(d) < 1e-5 ? 0.0 : p2 / (d * d);
  }
  term[e] = t;
}

// wpl2 = sum_irk coef_irk sum_{numin <= ie <= nomx} |mmatvv[irk][ie - mst][ie]|^2/3 kwfer[irk][ie]
__global__ void __launch_bounds__(256) pygw_hd_wpl2(const double2* __restrict__ mmatvv, const double* __restrict__ kwfer,
                                                   int ldE, const double* __restrict__ weight, int nirk, int nvbm,
                                                   int n_e, int numin, double c4, double* __restrict__ wpl2) {
  __shared__ double red[256];
  double acc = 0.0;
  const int nd = nvbm - numin;
  for (int e = threadIdx.x; e < nirk * std::max(nd, 0); e += blockDim.x) {
    const int irk = e / nd, ie = numin + e % nd;
    const double2* m = mmatvv + (((size_t)irk * n_e + (ie - numin)) * nvbm + ie) * 3;
    const double p2 = (m[0].x * m[0].x + m[0].y * m[0].y + m[1].x * m[1].x + m[1].y * m[1].y + m[2].x * m[2].x +
                       m[2].y * m[2].y) / 3.0;
    acc += c4 * weight[irk] * p2 * kwfer[(size_t)irk * ldE + ie];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *wpl2 = red[0];
}

// part[iom][ch] = sum over the irk of chunk ch of c4 w_irk sum_pair term kcw
// (one block per (frequency, chunk of irk): kcw's (pair) slab of a k and
// frequency is contiguous)
__global__ void __launch_bounds__(256) pygw_hd_head_part(const double* __restrict__ term,
                                                        const double* __restrict__ kcw, long n_pair, int nk_kcw,
                                                        const int* __restrict__ k_ind,
                                                        const double* __restrict__ weight, int nirk, int per,
                                                        double c4, double* __restrict__ part) {
  __shared__ double red[256];
  const int iom = blockIdx.x, ch = blockIdx.y;
  double acc = 0.0;
  for (int irk = ch * per; irk < std::min(nirk, (ch + 1) * per); irk++) {
    const int ik = k_ind[irk];
    if (ik < 0 || ik >= nk_kcw) {  // not a k of kcw: the head reads as NaN
      acc = __builtin_nan("");
      continue;
    }
    const double* kc = kcw + ((size_t)iom * nk_kcw + ik) * n_pair;
    const double* t = term + (size_t)irk * n_pair;
    double s = 0.0;
    for (long p = threadIdx.x; p < n_pair; p += blockDim.x) s += t[p] * kc[p];
    acc += c4 * weight[irk] * s;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(size_t)iom * gridDim.y + ch] = red[0];
}

// head[iom] = -sum_ch part[iom][ch] (in order) + the Drude term
__global__ void __launch_bounds__(256) pygw_hd_head_fin(const double* __restrict__ part, int nch,
                                                       const double* __restrict__ wpl2,
                                                       const double* __restrict__ omega,
                                                       const double* __restrict__ dUl, int nom_w, int nom,
                                                       double eta_head, double2* __restrict__ head) {
  const int iom = blockIdx.x * blockDim.x + threadIdx.x;
  if (iom >= nom) return;
  double h = 0.0;
  for (int c = 0; c < nch; c++) h -= part[(size_t)iom * nch + c];
  if (wpl2) {
    if (dUl) {
      for (int w = 0; w < nom_w; w++) {
        const double x = omega[w] + eta_head;
        h += *wpl2 / (x * x) * dUl[(size_t)w * nom + iom];
      }
    } else {
      const double x = omega[iom] + eta_head;
      h += *wpl2 / (x * x);
    }
  }
  head[iom] = double2{h, 0.0};
}

// pm[k][ie12] (calc_eps's q = 0 wing): sum_c p_c coefw/sqrt(3)/(e1 - e2)
__global__ void __launch_bounds__(256) pygw_hd_pm(const double2* __restrict__ mmatvv, const double2* __restrict__ mmatcv,
                                                 const double* __restrict__ Ebnd, int ldE,
                                                 const double* __restrict__ eig_core, const int* __restrict__ kii,
                                                 int nirk, int nk, int nvbm, int ncg, int ncg_p, int n_e, int numin,
                                                 double cw, double2* __restrict__ pm) {
  const long n_o = ncg_p + nvbm;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= nk * n_o * n_e) return;
  const int ie2 = (int)(e % n_e), io = (int)((e / n_e) % n_o), k = (int)(e / (n_o * n_e));
  const int irk = kii[k];
  if (irk < 0 || irk >= nirk) {  // not an irreducible k of mmatvv: the row reads as NaN
    pm[e] = double2{__builtin_nan(""), __builtin_nan("")};
    return;
  }
  const double e2 = Ebnd[(size_t)irk * ldE + numin + ie2];
  const double2* m;
  double e1;
  if (io < ncg_p) {
    m = mmatcv + (((size_t)irk * n_e + ie2) * ncg + io) * 3;
    e1 = eig_core[io];
  } else {
    m = mmatvv + (((size_t)irk * n_e + ie2) * nvbm + (io - ncg_p)) * 3;
    e1 = Ebnd[(size_t)irk * ldE + io - ncg_p];
  }
  const double d = e1 - e2;
  double2 v{0.0, 0.0};
  if (fabs(d) > 1e-10) {
    const double2 s{(m[0].x + m[1].x + m[2].x) * cw, (m[0].y + m[1].y + m[2].y) * cw};
    v = double2{s.x / d, s.y / d};
  }
  pm[e] = v;
}

static bool fill_layout(hd::Layout& L, int nat, const int* mult, const int* nLO_at, int nrows, int lmax, int nbmax,
                        int nbmax_g, int nlomax, int nlo, int lox) {
  using namespace hd;
  if (nat <= 0 || nat > kMaxNat || !mult || nrows < 0 || nrows > kMaxRows || (nrows > 0 && !nLO_at) || lmax <= 0 ||
      nbmax <= 0 || nlomax < 0 || nlo < 0 || lox < 0)
    return false;
  L = Layout{};
  L.nat = nat;
  L.lmax = lmax;
  L.ntnt = (lmax + 1) * (lmax + 1);
  L.nbmax = nbmax;
  L.nbmax_g = nbmax_g;
  L.nlomax = nlomax;
  L.nlo = nlo;
  L.lox = lox;
  L.nrows = nrows;
  for (int a = 0; a < nat; a++) {
    if (mult[a] <= 0 || L.ndf + mult[a] > kMaxNdf) return false;
    L.mult[a] = mult[a];
    for (int q = 0; q < mult[a]; q++) L.df_atom[L.ndf++] = a;
    for (int l = 0; l < nrows; l++) {
      const int n = nLO_at[l + nrows * a];
      if (n < 0) return false;
      // the gama slots and the local-orbital tables must hold every count
      // the loops reach: local orbitals at l < lmax are read at table row l
      // (pair (l, l+1)) and at l >= 1 at row l - 1 (pair (l-1, l))
      if (n > 0 && l <= lmax && (n > nlomax || n > nlo || (l < lmax && l >= lox) || (l >= 1 && l - 1 >= lox)))
        return false;
      L.nlo_at[l + nrows * a] = n;
    }
  }
  return true;
}

static size_t al256(size_t x) { return (x + 255) / 256 * 256; }

// workspace: Uc [nbw][nch], V [3][nbw][nch], the split products (32 partial
// planes of 6 n1 n2), O [nvk][nvk], the split overlap product (32 planes of
// nvk nbw) and T [nvk][nbw]
extern "C" int pygw_mmat_vv_workspace(int ndf, int lmax, int nlomax, int nvk, int nst, int nend, int mst, int mend,
                                      size_t* bytes) {
  if (ndf <= 0 || lmax <= 0 || nlomax < 0 || nvk < 0 || nend <= nst || mend <= mst || nst < 0 || mst < 0 || !bytes)
    return PYGW_E_ARG;
  const size_t nbw = std::max(nend, mend) - std::min(nst, mst), n1 = nend - nst, n2 = mend - mst;
  const size_t nch = (size_t)ndf * (lmax + 1) * (lmax + 1) * (2 + nlomax), nv = std::max(nvk, 1);
  *bytes = al256(nbw * nch * 16) + al256(3 * nbw * nch * 16) + al256(32 * 6 * n1 * n2 * 16) + al256(nv * nv * 16) +
           al256(32 * nv * nbw * 16) + al256(nv * nbw * 16);
  return PYGW_OK;
}

extern "C" int pygw_mmat_vv(int nat, const int* mult, const int* nLO_at, int nrows, int lmax, const void* alfa,
                            const void* beta, const void* gama, int nbmax, int nbmax_g, int nlomax,
                            const double* const* vv_tables, int nlo, int lox, const void* Aeig, long lda,
                            const double* kl, const int* indgk, int nvk, const void* ipwint, const int* gindex, int ng,
                            const int* i_g0, const int* n_g0, const double* k2cartes, int nst, int nend, int mst,
                            int mend, void* mmatvv, void* workspace, size_t ws_bytes, void* stream_) {
  using namespace hd;
  hipStream_t stream = (hipStream_t)stream_;
  Layout L;
  size_t need = 0;
  if (!fill_layout(L, nat, mult, nLO_at, nrows, lmax, nbmax, nbmax_g, nlomax, nlo, lox) || !alfa || !beta ||
      (nlomax > 0 && !gama) || !vv_tables || !Aeig || !kl || !mmatvv || nst < 0 || mst < 0 || nend > nbmax ||
      mend > nbmax || (nlomax > 0 && (nend > nbmax_g || mend > nbmax_g)) || !workspace ||
      pygw_mmat_vv_workspace(L.ndf, lmax, nlomax, nvk, nst, nend, mst, mend, &need) != PYGW_OK || ws_bytes < need)
    return PYGW_E_ARG;
  for (int i = 0; i < 9; i++)
    if (!vv_tables[i] && (i < 4 || nlo > 0)) return PYGW_E_ARG;
  if (nvk > 0 && (!indgk || !ipwint || !gindex || ng <= 0 || !i_g0 || !n_g0 || !k2cartes || lda < nvk ||
                  n_g0[0] < 0 || n_g0[1] < 0 || n_g0[2] < 0))
    return PYGW_E_ARG;
  const int n1 = nend - nst, n2 = mend - mst;
  const int b0 = std::min(nst, mst), nbw = std::max(nend, mend) - b0;
  const long nch = (long)L.ndf * L.ntnt * (2 + nlomax);
  const size_t nv = std::max(nvk, 1);
  const VvTables T{vv_tables[0], vv_tables[1], vv_tables[2], vv_tables[3], vv_tables[4],
                   vv_tables[5], vv_tables[6], vv_tables[7], vv_tables[8]};
  const Bands C{(const double2*)alfa, (const double2*)beta, (const double2*)gama, nbmax, nbmax_g, L.ntnt, nlomax};
  char* ws = (char*)workspace;
  double2* Uc = (double2*)ws;
  ws += al256((size_t)nbw * nch * 16);
  double2* V = (double2*)ws;
  ws += al256(3 * (size_t)nbw * nch * 16);
  double2* P = (double2*)ws;
  ws += al256(32 * 6 * (size_t)n1 * n2 * 16);
  double2* O = (double2*)ws;
  ws += al256(nv * nv * 16);
  double2* Tp = (double2*)ws;
  ws += al256(32 * nv * nbw * 16);
  double2* Tm = (double2*)ws;
  hipLaunchKernelGGL(pygw_hd_kapply, dim3(ceil_div((long)nbw * nch, 256)), dim3(256), 0, stream, L, T, C, b0, nbw,
                     nch, Uc, V);
  Job jobs[6];
  for (int c = 0; c < 3; c++) {
    jobs[c] = Job{Uc + (size_t)(nst - b0) * nch, V + ((size_t)c * nbw + (mst - b0)) * nch, nch, nch, 0, n1, n2, 0, 0};
    jobs[3 + c] = Job{Uc + (size_t)(mst - b0) * nch, V + ((size_t)c * nbw + (nst - b0)) * nch, nch, nch, 0, n2, n1,
                      0, 0};
  }
  const Jobs Jm = make_jobs(6, jobs, nch, 0);
  hipLaunchKernelGGL(pygw_hd_gemm, dim3(Jm.ntile, n_split(Jm)), dim3(256), 0, stream, Jm, P);
  hipLaunchKernelGGL(pygw_hd_mtcombine, dim3(ceil_div((long)n1 * n2 * 3, 256)), dim3(256), 0, stream,
                     (const double2*)P, n_split(Jm), Jm.plane, n1, n2, (double2*)mmatvv);
  if (nvk > 0) {
    hipLaunchKernelGGL(pygw_hd_olap, dim3(ceil_div((long)nvk * nvk, 256)), dim3(256), 0, stream, indgk, nvk, gindex,
                       ng, (const double2*)ipwint, i_g0, n_g0[0], n_g0[1], n_g0[2], O);
    const Job jt{O, (const double2*)Aeig + (size_t)b0 * lda, (long)nvk, lda, 0, nvk, nbw, 0, 0};
    const Jobs Jt = make_jobs(1, &jt, nvk, 1);
    hipLaunchKernelGGL(pygw_hd_gemm, dim3(Jt.ntile, n_split(Jt)), dim3(256), 0, stream, Jt, Tp);
    hipLaunchKernelGGL(pygw_hd_sum, dim3(ceil_div(Jt.plane, 256)), dim3(256), 0, stream, (const double2*)Tp,
                       n_split(Jt), Jt.plane, Tm);
    KpG kp;
    for (int i = 0; i < 9; i++) kp.k2c[i] = k2cartes[i];
    for (int c = 0; c < 3; c++) kp.kcar[c] = k2cartes[c] * kl[0] + k2cartes[c + 3] * kl[1] + k2cartes[c + 6] * kl[2];
    hipLaunchKernelGGL(pygw_hd_is_p, dim3(ceil_div((long)n1 * n2, 4)), dim3(256), 0, stream, (const double2*)Aeig,
                       lda, b0, nbw, (const double2*)Tm, indgk, nvk, gindex, ng, kp, nst, n1, mst, n2,
                       (double2*)mmatvv);
  }
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

extern "C" int pygw_mmat_cv(int nat, const int* mult, const int* nLO_at, int nrows, int lmax, const int* core_off,
                            const void* alfa, const void* beta, const void* gama, int nbmax, int nbmax_g, int nlomax,
                            const double* const* cv_tables, int nloat, int nct, const int* corind, int ncg, int mst,
                            int mend, int exact, void* mmatcv, void* stream) {
  using namespace hd;
  Layout L;
  if (!fill_layout(L, nat, mult, nLO_at, nrows, lmax, nbmax, nbmax_g, nlomax, 1 << 20, 1 << 20) || !core_off ||
      !alfa || !beta || (nlomax > 0 && !gama) || !cv_tables || nct <= 0 || nloat < 0 || !corind || ncg <= 0 ||
      mst < 0 || mend <= mst || mend > nbmax || (nlomax > 0 && mend > nbmax_g) || !mmatcv)
    return PYGW_E_ARG;
  for (int i = 0; i < 6; i++)
    if (!cv_tables[i] && (i < 4 || nloat > 0)) return PYGW_E_ARG;
  for (int a = 0; a < nat; a++) {
    // local orbitals at l +- 1 of a core state read iulol_ucl / iucl_ulol up to nLO_at
    for (int l = 0; l < nrows; l++)
      if (l <= lmax && L.nlo_at[l + nrows * a] > nloat) return PYGW_E_ARG;
    if (core_off[a] < 0 || core_off[a] > nct) return PYGW_E_ARG;
    L.core_off[a] = core_off[a];
  }
  const CvTables T{cv_tables[0], cv_tables[1], cv_tables[2], cv_tables[3], cv_tables[4], cv_tables[5], nloat, nct};
  const Bands C{(const double2*)alfa, (const double2*)beta, (const double2*)gama, nbmax, nbmax_g, L.ntnt, nlomax};
  const int n2 = mend - mst;
  hipLaunchKernelGGL(pygw_hd_cv, dim3(ceil_div((long)ncg * n2, 256)), dim3(256), 0, (hipStream_t)stream, L, T, C,
                     corind, ncg, mst, n2, exact ? 1 : 0, (double2*)mmatcv);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

namespace hd {
constexpr int kHeadChunks = 16;  // blocks per frequency in the head sums
constexpr int kMaxNom = 4096;
}  // namespace hd

extern "C" int pygw_calc_head_workspace(int nirk, int n_o, int n_e, size_t* bytes) {
  if (nirk <= 0 || n_o <= 0 || n_e <= 0 || !bytes) return PYGW_E_ARG;
  *bytes = 256 + (size_t)hd::kMaxNom * hd::kHeadChunks * sizeof(double) + (size_t)nirk * n_o * n_e * sizeof(double);
  return PYGW_OK;
}

extern "C" int pygw_calc_head(const void* mmatvv, const void* mmatcv, const double* Ebnd, int ldE,
                              const double* eig_core, int ncg, const double* kcw, int nk_kcw, const int* k_ind,
                              const double* weight, int nirk, int nomx, int numin, int nbmaxpol, int ncg_p, int nom,
                              double vol, int metallic, const double* kwfer, int iop_drude, double eta_head,
                              const double* omega, const double* dUl, int nom_w, void* head, void* workspace,
                              size_t ws_bytes, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  const int nvbm = nomx + 1, n_e = nbmaxpol - numin, n_o = ncg_p + nvbm;
  size_t need = 0;
  const bool drude = metallic && iop_drude == 1;
  if (!mmatvv || (ncg_p > 0 && (!mmatcv || !eig_core || ncg < ncg_p)) || !Ebnd || !kcw || !k_ind || !weight ||
      nomx < 0 || numin < 0 || n_e <= 0 || ldE < nbmaxpol || nom <= 0 || nom > hd::kMaxNom || nk_kcw <= 0 ||
      !(vol > 0.0) || !head ||
      ncg_p < 0 || pygw_calc_head_workspace(nirk, n_o, n_e, &need) != PYGW_OK || !workspace || ws_bytes < need ||
      (drude && (!kwfer || !omega || (dUl ? nom_w <= 0 : nom_w != nom))))
    return PYGW_E_ARG;
  const double c4 = 4.0 * 3.14159265358979323846 * 2.0 / vol;  // 4 pi fspin / Vol, fspin = 2
  double* wpl2 = (double*)workspace;
  double* part = (double*)((char*)workspace + 256);
  double* term = part + (size_t)hd::kMaxNom * hd::kHeadChunks;
  hipLaunchKernelGGL(pygw_hd_terms, dim3(ceil_div((long)nirk * n_o * n_e, 256)), dim3(256), 0, stream,
                     (const double2*)mmatvv, (const double2*)mmatcv, Ebnd, ldE, eig_core, nirk, nvbm, ncg, ncg_p, n_e,
                     numin, term);
  if (drude)
    hipLaunchKernelGGL(pygw_hd_wpl2, dim3(1), dim3(256), 0, stream, (const double2*)mmatvv, kwfer, ldE, weight, nirk,
                       nvbm, n_e, numin, c4, wpl2);
  const int nch = std::min(hd::kHeadChunks, nirk), per = ceil_div(nirk, nch);
  hipLaunchKernelGGL(pygw_hd_head_part, dim3(nom, nch), dim3(256), 0, stream, term, kcw, (long)n_o * n_e, nk_kcw,
                     k_ind, weight, nirk, per, c4, part);
  hipLaunchKernelGGL(pygw_hd_head_fin, dim3(ceil_div(nom, 256)), dim3(256), 0, stream, (const double*)part, nch,
                     drude ? (const double*)wpl2 : nullptr, omega, dUl, nom_w, nom, eta_head, (double2*)head);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

extern "C" int pygw_wing_pm(const void* mmatvv, const void* mmatcv, const double* Ebnd, int ldE,
                            const double* eig_core, int ncg, int nirk, const int* kii, int nk, int nomx, int numin,
                            int nbmaxpol, int ncg_p, double vol, void* pm, void* stream) {
  const int nvbm = nomx + 1, n_e = nbmaxpol - numin;
  if (!mmatvv || (ncg_p > 0 && (!mmatcv || !eig_core || ncg < ncg_p)) || !Ebnd || ldE < nbmaxpol || !kii ||
      nirk <= 0 || nk <= 0 || nomx < 0 || numin < 0 || n_e <= 0 || ncg_p < 0 || !(vol > 0.0) || !pm)
    return PYGW_E_ARG;
  const double cw = 2.0 * sqrt(3.14159265358979323846 / vol) / sqrt(3.0);
  hipLaunchKernelGGL(pygw_hd_pm, dim3(ceil_div((long)nk * (ncg_p + nvbm) * n_e, 256)), dim3(256), 0,
                     (hipStream_t)stream, (const double2*)mmatvv, (const double2*)mmatcv, Ebnd, ldE, eig_core, kii,
                     nirk, nk, nvbm, ncg, ncg_p, n_e, numin, cw, (double2*)pm);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}
