// The muffin-tin and core parts of the product-basis matrix elements on
// gfx950 (include/pygw_hip.h):
//
//   pygw_minm_mt: calc_minm_MT (for_vxcnn.f90:839-1028, pygw.py:2655),
//     mmat[ie1][ie2][im] = sum_c N_im[ie1][c] P_idf[c][ie2], with c the
//     radial channels of the band at k - q (u, udot of every (l2, m2), then the
//     local orbitals in (l2, m2, ilo2) order), P the coefficients alfp / betp /
//     gamp of those channels, and
//       N_im[ie1][c] = e^{-2 pi i ql.pos} sum_{l1} <Y_l1,m2+mb | Y_l2m2 Y_lbmb>
//                      sum_{x1} s3r(irm, x1(l1), x2(c)) A_x1[ie1][l1, m2 + mb]
//     (x1 over u, udot and the local orbitals of l1; A = alfa, beta, gama).
//     N is built by one thread per entry (pygw_mmt_build_n) in the order of
//     the reference's sums; the product with P is an FP64-MFMA GEMM
//     (pygw_mmt_gemm) whose tiles cover 64 product functions of one atom for
//     one band ie1, so that each row of mmat is written in runs of
//     consecutive im.
//   pygw_minc / pygw_minc2: calc_minc / calc_minc2 (for_vxcnn.f90:1136-1397,
//     pygw.py:2677-2680), the products of a valence band with a core state:
//     one thread per entry, the l1 sum in the reference's order.
//
// The atom / product-function / channel structure is a plan of int32
// (pygw_minm_mt_plan, host), passed both in host memory (launch geometry)
// and in device memory (the kernels' tables).
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"
#include "../../include/pygw_hip.h"

namespace mm {
// the value of pi calc_minm_MT / calc_minc carry (for_vxcnn.f90:877, 1160)
constexpr double kPiF = 3.14159265358979;
// plan header (ints)
enum {
  H_NAT, H_NDF, H_LMAX, H_NLM, H_CMAX, H_LOCT, H_NTILE, H_OFF_DF, H_OFF_IM, H_OFF_CH, H_OFF_TY, H_OFF_TILE, H_SIZE
};
constexpr int DF_W = 4;   // per atom: type, im0, n_im, C
constexpr int IM_W = 4;   // per product function: idf, irm, lb, mb
constexpr int CH_W = 4;   // per (type, channel): kind (0 u, 1 udot, 2 lo), l, m, s3r index
constexpr int TY_W = 2;   // per type: ncore, C; then per l: LO count, first LO s3r index
constexpr int TILE_W = 2; // per GEMM tile: idf, first im
constexpr int TR = 64, TC = 64, KC = 16;
}  // namespace mm

// <Y_l3,m1+m2 | Y_l1m1 Y_l2m2> from the table of cmp_all_Gaunt (getcgcoef,
// gaunt.f90:36-101); 0 outside the triangle or the table.
__device__ inline double getcgcoef(int l1, int l2, int l3, int m1, int m2, const double* __restrict__ cg, int ncg) {
  const int par = abs(l1 + l2 - l3) % 2;
  if (!(abs(m1 + m2) <= l3 && abs(l1 - l2) <= l3 && l1 + l2 >= l3)) return 0.0;
  int j1, mj1, j2, mj2;
  if (l1 < l2) {
    j1 = l2; mj1 = m2; j2 = l1; mj2 = m1;
  } else {
    j1 = l1; mj1 = m1; j2 = l2; mj2 = m2;
  }
  double fact = 1.0;
  if (mj2 < 0) {
    mj2 = -mj2;
    mj1 = -mj1;
    fact = -2.0 * par + 1.0;
  }
  const long ind1 = (long)(16 * j1 * j1 - 3 * j1 - 3) * (j1 + 2) * (j1 + 1) * j1 / 60;
  const long ind2 = (long)j1 * j2 * (j2 + 1) * (4 * j2 - 1) / 3;
  const long ind3 = (long)j2 * (j2 - 1) * (4 * j2 + 7) / 6;
  const long ind4 = (long)(2 * j1 + 1) * (j2 + 1) * (l3 - j1 + j2);
  const long ing = ind1 + ind2 + ind3 + ind4 + (j2 + 1) * (mj1 + j1) + mj2 + j2 + 1;
  return ing >= 1 && ing <= ncg ? fact * cg[ing - 1] : __builtin_nan("");
}

__device__ inline double2 cmulc(double2 a, double2 b) { return double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ inline double2 cadd(double2 a, double2 b) { return double2{a.x + b.x, a.y + b.y}; }
__device__ inline double2 rmul(double2 a, double r) { return double2{a.x * r, a.y * r}; }

// e^{sgn 2 pi i v.pos_idf}
__device__ inline double2 atom_phase(const double* __restrict__ v, const double* __restrict__ pos, int idf, double sgn) {
  const double arg = sgn * 2.0 * mm::kPiF * (v[0] * pos[3 * idf] + v[1] * pos[3 * idf + 1] + v[2] * pos[3 * idf + 2]);
  double s, c;
  sincos(arg, &s, &c);
  return double2{c, s};
}

// P[(idf * cmax + c) * n2 + ie2]: the coefficient of band mstart + ie2 in
// channel c of atom idf (alfp, betp or gamp).  One thread per entry.
__global__ void __launch_bounds__(256) pygw_mmt_build_p(const int* __restrict__ plan, int mstart, int n2,
                                                        const double2* __restrict__ alfp,
                                                        const double2* __restrict__ betp,
                                                        const double2* __restrict__ gamp, int mbmax, int mbmax_g,
                                                        int nlomax, double2* __restrict__ P) {
  using namespace mm;
  const int ndf = plan[H_NDF], nlm = plan[H_NLM], cmax = plan[H_CMAX];
  const long e = blockIdx.x * 256L + threadIdx.x;
  if (e >= (long)ndf * cmax * n2) return;
  const int ie2 = (int)(e % n2), c = (int)((e / n2) % cmax), idf = (int)(e / ((long)n2 * cmax));
  const int* df = plan + plan[H_OFF_DF] + DF_W * idf;
  const int ty = df[0];
  double2 v{0.0, 0.0};
  if (c < df[3]) {
    const int* ch = plan + plan[H_OFF_CH] + CH_W * (ty * cmax + c);
    const int kind = ch[0], l = ch[1], m = ch[2];
    const int lm = l * l + l + m;
    const long ie = ie2 + mstart;
    if (kind == 0)
      v = alfp[ie + (long)mbmax * (lm + (long)nlm * idf)];
    else if (kind == 1)
      v = betp[ie + (long)mbmax * (lm + (long)nlm * idf)];
    else {
      // the local orbital's number within (l, m): its s3r index minus the first of l
      const int* tyr = plan + plan[H_OFF_TY] + ty * (TY_W + 2 * (plan[H_LMAX] + 1));
      const int ilo = ch[3] - tyr[TY_W + 2 * l + 1];
      v = gamp[ie + (long)mbmax_g * (ilo + (long)nlomax * (lm + (long)nlm * idf))];
    }
  }
  P[e] = v;
}

// NT[(ie1 * cmax + c) * loct + im] = N_im[ie1][c] for c < C of im's atom
// type.  One thread per entry, im fastest.
__global__ void __launch_bounds__(256) pygw_mmt_build_n(const int* __restrict__ plan, const double* __restrict__ ql,
                                                        const double* __restrict__ pos, int nstart, int n1,
                                                        const double2* __restrict__ alfa,
                                                        const double2* __restrict__ beta,
                                                        const double2* __restrict__ gama, int nbmax, int nbmax_g,
                                                        int nlomax, const double* __restrict__ s3r, int n_mix,
                                                        int n_fnc, const double* __restrict__ cg, int ncg,
                                                        double2* __restrict__ NT) {
  using namespace mm;
  const int loct = plan[H_LOCT], cmax = plan[H_CMAX], lmax = plan[H_LMAX], nlm = plan[H_NLM];
  const long e = blockIdx.x * 256L + threadIdx.x;
  if (e >= (long)n1 * cmax * loct) return;
  const int im = (int)(e % loct), c = (int)((e / loct) % cmax), ie1 = (int)(e / ((long)loct * cmax));
  const int* imr = plan + plan[H_OFF_IM] + IM_W * im;
  const int idf = imr[0], irm = imr[1], lb = imr[2], mb = imr[3];
  const int* df = plan + plan[H_OFF_DF] + DF_W * idf;
  const int ty = df[0];
  if (c >= df[3]) return;
  const int* tyr = plan + plan[H_OFF_TY] + ty * (TY_W + 2 * (lmax + 1));
  const int ncore = tyr[0];
  const int* ch = plan + plan[H_OFF_CH] + CH_W * (ty * cmax + c);
  const int kind2 = ch[0], l2 = ch[1], m2 = ch[2], x2 = ch[3];
  const double2 phs = atom_phase(ql, pos, idf, -1.0);
  const int m1 = m2 + mb;
  // s3r(irm, x1, x2, ty), Fortran (n_mix, n_fnc, n_fnc, nat)
  const double* s3 = s3r + irm + (long)n_mix * (long)n_fnc * ((long)x2 + (long)n_fnc * ty);
  const long ie = ie1 + nstart;
  double2 s{0.0, 0.0};
  for (int l1 = abs(l2 - lb); l1 <= min(l2 + lb, lmax); l1++) {
    if (abs(m1) > l1) continue;
    const double ang = getcgcoef(l2, lb, l1, m2, mb, cg, ncg);
    if (fabs(ang) < 1e-8) continue;
    const double2 pa = rmul(phs, ang);
    const int lm1 = l1 * l1 + l1 + m1;
    const int la1 = ncore + l1, lb1 = la1 + lmax + 1;
    const int nlo1 = tyr[TY_W + 2 * l1], lc10 = tyr[TY_W + 2 * l1 + 1];
    const double2 a1 = alfa[ie + (long)nbmax * (lm1 + (long)nlm * idf)];
    const double2 b1 = beta[ie + (long)nbmax * (lm1 + (long)nlm * idf)];
    const double2 aa = rmul(pa, s3[(long)n_mix * la1]), ab = rmul(pa, s3[(long)n_mix * lb1]);
    if (kind2 < 2) {
      // na / nb: += aa alfa + ab beta, then += ac(ilo1) gama per local orbital of l1
      s = cadd(cadd(s, cmulc(aa, a1)), cmulc(ab, b1));
      for (int ilo1 = 0; ilo1 < nlo1; ilo1++) {
        const double2 g1 = gama[ie + (long)nbmax_g * (ilo1 + (long)nlomax * (lm1 + (long)nlm * idf))];
        s = cadd(s, cmulc(rmul(pa, s3[(long)n_mix * (lc10 + ilo1)]), g1));
      }
    } else {
      // nc: += sum_ilo1 cc gama + ca alfa + cb beta
      double2 ccsm{0.0, 0.0};
      for (int ilo1 = 0; ilo1 < nlo1; ilo1++) {
        const double2 g1 = gama[ie + (long)nbmax_g * (ilo1 + (long)nlomax * (lm1 + (long)nlm * idf))];
        ccsm = cadd(ccsm, cmulc(rmul(pa, s3[(long)n_mix * (lc10 + ilo1)]), g1));
      }
      s = cadd(cadd(cadd(s, ccsm), cmulc(aa, a1)), cmulc(ab, b1));
    }
  }
  NT[e] = s;
}

// mmat[(ie1 n2 + ie2) ld + col0 + im] = sum_{c < C} P[idf][c][ie2] NT[ie1][c][im]
// for the 64 product functions of one tile (one atom, blockIdx.x), one band
// ie1 (blockIdx.y) and 64 bands ie2 (blockIdx.z): four waves of 32 x 32 (2 x 2
// FP64 MFMA 16 x 16 x 4 tiles, real and imaginary parts).
__global__ void __launch_bounds__(256) pygw_mmt_gemm(const int* __restrict__ plan, int n1, int n2,
                                                     const double2* __restrict__ P, const double2* __restrict__ NT,
                                                     double2* __restrict__ mmat, long ld, int col0) {
  using namespace mm;
  typedef double d4 __attribute__((ext_vector_type(4)));
  __shared__ double2 As[TR][KC + 1];   // P^T tile: rows ie2, columns c
  __shared__ double2 Bs[KC][TC + 1];   // NT tile: rows c, columns im
  const int cmax = plan[H_CMAX], loct = plan[H_LOCT];
  const int* tile = plan + plan[H_OFF_TILE] + TILE_W * blockIdx.x;
  const int idf = tile[0], im0 = tile[1];
  const int* df = plan + plan[H_OFF_DF] + DF_W * idf;
  const int im_end = df[1] + df[2], C = df[3];
  const int ie1 = blockIdx.y, r0 = blockIdx.z * TR;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1, li = lane & 15, lk = lane >> 4;
  const double2* Pa = P + (long)idf * cmax * n2;
  const double2* Nb = NT + (long)ie1 * cmax * loct;
  d4 aR[2][2], aI[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      aR[a][b] = d4{0, 0, 0, 0};
      aI[a][b] = d4{0, 0, 0, 0};
    }
  for (int k0 = 0; k0 < C; k0 += KC) {
    for (int e = tid; e < TR * KC; e += 256) {
      const int r = e % TR, c = e / TR;
      const int ie2 = r0 + r, k = k0 + c;
      As[r][c] = (ie2 < n2 && k < C) ? Pa[(long)k * n2 + ie2] : double2{0.0, 0.0};
    }
    for (int e = tid; e < KC * TC; e += 256) {
      const int j = e % TC, c = e / TC;
      const int im = im0 + j, k = k0 + c;
      Bs[c][j] = (im < im_end && k < C) ? Nb[(long)k * loct + im] : double2{0.0, 0.0};
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
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int ie2 = r0 + 32 * wr + 16 * a + 4 * v + lk;
      if (ie2 >= n2) continue;
#pragma unroll
      for (int b = 0; b < 2; b++) {
        const int im = im0 + 32 * wc + 16 * b + li;
        if (im < im_end) mmat[((long)ie1 * n2 + ie2) * ld + col0 + im] = double2{aR[a][b][v], aI[a][b][v]};
      }
    }
}

// calc_minc (sgn -1, m1 = mc + mb, <Y_l1,m1 | Y_lc,mc Y_lb,mb>, band index
// ie1 at k) and calc_minc2 (sgn +1, m1 = mc - mb, <Y_lc,mc | Y_l1,m1 Y_lb,mb>,
// band ie2 at k - q): out[(row) ld + col0 + im] with row = ie * nc + icg
// (first) or icg * nb + ie (second).  One thread per entry, im fastest; 0
// unless the core state sits on the product function's atom.
template <bool SECOND>
__global__ void __launch_bounds__(256) pygw_minc_k(const int* __restrict__ plan, const double* __restrict__ kv,
                                                   const double* __restrict__ pos, int bstart, int nb, int cstart,
                                                   int nc, const double2* __restrict__ alfa,
                                                   const double2* __restrict__ beta, const double2* __restrict__ gama,
                                                   int nbmax, int nbmax_g, int nlomax, const double* __restrict__ s3r,
                                                   int n_mix, int n_fnc, const int* __restrict__ corind, int ncg_rows,
                                                   const double* __restrict__ cg, int ncg, double2* __restrict__ out,
                                                   long ld, int col0) {
  using namespace mm;
  const int loct = plan[H_LOCT], lmax = plan[H_LMAX], nlm = plan[H_NLM];
  const long e = blockIdx.x * 256L + threadIdx.x;
  if (e >= (long)nb * nc * loct) return;
  const int im = (int)(e % loct);
  const long rowi = e / loct;
  const int ib = SECOND ? (int)(rowi % nb) : (int)(rowi / nc);
  const int ic = SECOND ? (int)(rowi / nb) : (int)(rowi % nc);
  const int* imr = plan + plan[H_OFF_IM] + IM_W * im;
  const int idf = imr[0], irm = imr[1], lb = imr[2], mb = imr[3];
  const int icg = cstart + ic;
  // corind(ncg, 5) Fortran: atom type, atom, core index, l, m (0-based)
  const int cty = corind[icg], cdf = corind[icg + ncg_rows], cix = corind[icg + 2 * ncg_rows];
  const int lc = corind[icg + 3 * ncg_rows], mc = corind[icg + 4 * ncg_rows];
  double2 s{0.0, 0.0};
  if (cdf == idf) {
    const int* tyr = plan + plan[H_OFF_TY] + cty * (TY_W + 2 * (lmax + 1));
    const int ncore = tyr[0];
    const double2 phs = atom_phase(kv, pos, idf, SECOND ? 1.0 : -1.0);
    const double* s3 = s3r + irm + (long)n_mix * ((long)cix + (long)n_fnc * (long)n_fnc * cty);
    const long ie = ib + bstart;
    for (int l1 = abs(lb - lc); l1 <= min(lb + lc, lmax); l1++) {
      const int m1 = SECOND ? mc - mb : mc + mb;
      const double ang = SECOND ? getcgcoef(l1, lb, lc, m1, mb, cg, ncg) : getcgcoef(lc, lb, l1, mc, mb, cg, ncg);
      if (abs(m1) > l1 || fabs(ang) < 1e-8) continue;
      const double2 pa = rmul(phs, ang);
      const int lm1 = l1 * l1 + l1 + m1;
      const int la1 = ncore + l1, lb1 = la1 + lmax + 1;
      const int nlo1 = tyr[TY_W + 2 * l1], lc10 = tyr[TY_W + 2 * l1 + 1];
      // s3r(irm, ic, x1, ty)
      const long sx = (long)n_mix * n_fnc;
      double2 csm = cadd(cmulc(rmul(pa, s3[sx * la1]), alfa[ie + (long)nbmax * (lm1 + (long)nlm * idf)]),
                         cmulc(rmul(pa, s3[sx * lb1]), beta[ie + (long)nbmax * (lm1 + (long)nlm * idf)]));
      for (int ilo1 = 0; ilo1 < nlo1; ilo1++)
        csm = cadd(csm, cmulc(rmul(pa, s3[sx * (lc10 + ilo1)]),
                              gama[ie + (long)nbmax_g * (ilo1 + (long)nlomax * (lm1 + (long)nlm * idf))]));
      s = cadd(s, csm);
    }
  }
  out[rowi * ld + col0 + im] = s;
}

// ---------------------------------------------------------------------------
extern "C" int pygw_minm_mt_plan(int nat, const int* mult, const int* nmix, const int* big_l, int max_nmix,
                                 const int* nLO_at, int lomax, const int* ncore, int lmax, int* plan, int* plan_ints) {
  using namespace mm;
  if (nat <= 0 || !mult || !nmix || !big_l || max_nmix <= 0 || !nLO_at || lomax < -1 || !ncore || lmax < 0 ||
      lmax > 16 || !plan_ints)
    return PYGW_E_ARG;
  const int nlm = (lmax + 1) * (lmax + 1);
  int ndf = 0, loct = 0, cmax = 0, ntile = 0;
  std::vector<int> C(nat), nim(nat);
  for (int a = 0; a < nat; a++) {
    if (mult[a] <= 0 || nmix[a] < 0 || nmix[a] > max_nmix || ncore[a] < 0) return PYGW_E_ARG;
    int ncl = 0;
    for (int l = 0; l <= std::min(lomax, lmax); l++) {
      if (nLO_at[l + (lomax + 1) * a] < 0) return PYGW_E_ARG;
      ncl += (2 * l + 1) * nLO_at[l + (lomax + 1) * a];
    }
    C[a] = 2 * nlm + ncl;
    cmax = std::max(cmax, C[a]);
    nim[a] = 0;
    for (int f = 0; f < nmix[a]; f++) {
      const int L = big_l[f + max_nmix * a];
      if (L < 0) return PYGW_E_ARG;
      nim[a] += 2 * L + 1;
    }
    ndf += mult[a];
    loct += mult[a] * nim[a];
    ntile += mult[a] * ((nim[a] + TR - 1) / TR);
  }
  const int off_df = H_SIZE, off_im = off_df + DF_W * ndf, off_ch = off_im + IM_W * loct;
  const int off_ty = off_ch + CH_W * nat * cmax, off_tile = off_ty + nat * (TY_W + 2 * (lmax + 1));
  const int total = off_tile + TILE_W * ntile;
  if (!plan) {
    *plan_ints = total;
    return PYGW_OK;
  }
  if (*plan_ints < total) return PYGW_E_ARG;
  std::fill(plan, plan + total, 0);
  int* h = plan;
  h[H_NAT] = nat;
  h[H_NDF] = ndf;
  h[H_LMAX] = lmax;
  h[H_NLM] = nlm;
  h[H_CMAX] = cmax;
  h[H_LOCT] = loct;
  h[H_NTILE] = ntile;
  h[H_OFF_DF] = off_df;
  h[H_OFF_IM] = off_im;
  h[H_OFF_CH] = off_ch;
  h[H_OFF_TY] = off_ty;
  h[H_OFF_TILE] = off_tile;
  int idf = 0, im = 0, t = 0;
  for (int a = 0; a < nat; a++) {
    // per type: ncore, C, then per l: LO count and the 0-based s3r index of its first LO
    int* ty = plan + off_ty + a * (TY_W + 2 * (lmax + 1));
    ty[0] = ncore[a];
    ty[1] = C[a];
    int lo0 = ncore[a] + 2 * (lmax + 1);
    for (int l = 0; l <= lmax; l++) {
      const int n = l <= lomax ? nLO_at[l + (lomax + 1) * a] : 0;
      ty[TY_W + 2 * l] = n;
      ty[TY_W + 2 * l + 1] = lo0;
      lo0 += n;
    }
    // channels: u (l, m), udot (l, m), then the local orbitals (l, m, ilo)
    int* ch = plan + off_ch + CH_W * a * cmax;
    int c = 0;
    for (int kind = 0; kind < 2; kind++)
      for (int l = 0; l <= lmax; l++)
        for (int m = -l; m <= l; m++, c++) {
          ch[CH_W * c] = kind;
          ch[CH_W * c + 1] = l;
          ch[CH_W * c + 2] = m;
          ch[CH_W * c + 3] = ncore[a] + l + kind * (lmax + 1);
        }
    for (int l = 0; l <= lmax; l++)
      for (int m = -l; m <= l; m++)
        for (int ilo = 0; ilo < ty[TY_W + 2 * l]; ilo++, c++) {
          ch[CH_W * c] = 2;
          ch[CH_W * c + 1] = l;
          ch[CH_W * c + 2] = m;
          ch[CH_W * c + 3] = ty[TY_W + 2 * l + 1] + ilo;
        }
    for (int q = 0; q < mult[a]; q++, idf++) {
      int* df = plan + off_df + DF_W * idf;
      df[0] = a;
      df[1] = im;
      df[2] = nim[a];
      df[3] = C[a];
      for (int j = 0; j < nim[a]; j += TR, t++) {
        plan[off_tile + TILE_W * t] = idf;
        plan[off_tile + TILE_W * t + 1] = im + j;
      }
      for (int f = 0; f < nmix[a]; f++) {
        const int L = big_l[f + max_nmix * a];
        for (int m = -L; m <= L; m++, im++) {
          int* r = plan + off_im + IM_W * im;
          r[0] = idf;
          r[1] = f;
          r[2] = L;
          r[3] = m;
        }
      }
    }
  }
  *plan_ints = total;
  return PYGW_OK;
}

static bool plan_ok(const int* p) {
  using namespace mm;
  return p && p[H_NAT] > 0 && p[H_NDF] > 0 && p[H_LOCT] > 0 && p[H_CMAX] > 0 && p[H_NLM] == (p[H_LMAX] + 1) * (p[H_LMAX] + 1);
}

// s3r (n_mix, n_fnc, n_fnc, nat) holds every product function and radial
// channel of the plan
static bool plan_fits_s3r(const int* p, int n_mix, int n_fnc) {
  using namespace mm;
  for (int a = 0; a < p[H_NAT]; a++) {
    const int* ty = p + p[H_OFF_TY] + a * (TY_W + 2 * (p[H_LMAX] + 1));
    int need = ty[0] + 2 * (p[H_LMAX] + 1);
    for (int l = 0; l <= p[H_LMAX]; l++) need += ty[TY_W + 2 * l];
    if (need > n_fnc) return false;
  }
  for (int im = 0; im < p[H_LOCT]; im++)
    if (p[p[H_OFF_IM] + IM_W * im + 1] >= n_mix) return false;
  return true;
}

// the largest number of local orbitals of one l in the plan (gama's second extent must hold it)
static int plan_max_lo(const int* p) {
  using namespace mm;
  int n = 0;
  for (int a = 0; a < p[H_NAT]; a++) {
    const int* ty = p + p[H_OFF_TY] + a * (TY_W + 2 * (p[H_LMAX] + 1));
    for (int l = 0; l <= p[H_LMAX]; l++) n = std::max(n, ty[TY_W + 2 * l]);
  }
  return n;
}

extern "C" int pygw_minm_mt_workspace(const int* plan_host, int n1, int n2, size_t* bytes) {
  using namespace mm;
  if (!plan_ok(plan_host) || n1 <= 0 || n2 <= 0 || !bytes) return PYGW_E_ARG;
  const size_t nt = (size_t)n1 * plan_host[H_CMAX] * plan_host[H_LOCT] * sizeof(double2);
  *bytes = (nt + 255) / 256 * 256 + (size_t)plan_host[H_NDF] * plan_host[H_CMAX] * n2 * sizeof(double2);
  return PYGW_OK;
}

extern "C" int pygw_minm_mt(const int* plan_host, const int* plan_dev, const double* ql, int nstart, int nend,
                            int mstart, int mend, const void* alfa, const void* beta, const void* gama, int nbmax,
                            int nbmax_g, const void* alfp, const void* betp, const void* gamp, int mbmax, int mbmax_g,
                            int nlomax, const double* s3r, int n_mix, int n_fnc, const double* pos,
                            const double* cgcoef, int n_cg, void* mmat, long ld_mmat, int col0, void* workspace,
                            size_t ws_bytes, void* stream_) {
  using namespace mm;
  hipStream_t stream = (hipStream_t)stream_;
  const int n1 = nend - nstart, n2 = mend - mstart;
  size_t need = 0;
  if (!plan_dev || !ql || !alfa || !beta || !alfp || !betp || !s3r || !pos || !cgcoef || !mmat || nstart < 0 ||
      mstart < 0 || nend > nbmax || mend > mbmax || n_mix <= 0 || n_fnc <= 0 || n_cg <= 0 || col0 < 0 ||
      pygw_minm_mt_workspace(plan_host, n1, n2, &need) != PYGW_OK || ld_mmat < (long)col0 + plan_host[H_LOCT] ||
      nlomax < plan_max_lo(plan_host) || (nlomax > 0 && (!gama || !gamp || nend > nbmax_g || mend > mbmax_g)) ||
      !plan_fits_s3r(plan_host, n_mix, n_fnc))
    return PYGW_E_ARG;
  char* ws = (char*)workspace;
  bool own = false;
  if (!ws) {
    HIP_TRY(hipMallocAsync((void**)&ws, need, stream));
    own = true;
  } else if (ws_bytes < need) {
    return PYGW_E_ARG;
  }
  const int cmax = plan_host[H_CMAX], loct = plan_host[H_LOCT], ndf = plan_host[H_NDF];
  const size_t nt_bytes = ((size_t)n1 * cmax * loct * sizeof(double2) + 255) / 256 * 256;
  double2* NT = (double2*)ws;
  double2* P = (double2*)(ws + nt_bytes);
  hipLaunchKernelGGL(pygw_mmt_build_p, dim3(ceil_div((long)ndf * cmax * n2, 256)), dim3(256), 0, stream, plan_dev,
                     mstart, n2, (const double2*)alfp, (const double2*)betp, (const double2*)gamp, mbmax, mbmax_g,
                     nlomax, P);
  hipLaunchKernelGGL(pygw_mmt_build_n, dim3(ceil_div((long)n1 * cmax * loct, 256)), dim3(256), 0, stream, plan_dev, ql,
                     pos, nstart, n1, (const double2*)alfa, (const double2*)beta, (const double2*)gama, nbmax, nbmax_g,
                     nlomax, s3r, n_mix, n_fnc, cgcoef, n_cg, NT);
  hipLaunchKernelGGL(pygw_mmt_gemm, dim3(plan_host[H_NTILE], n1, ceil_div(n2, TR)), dim3(256), 0, stream, plan_dev,
                     n1, n2, (const double2*)P, (const double2*)NT, (double2*)mmat, ld_mmat, col0);
  int rc = hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
  if (own && hipFreeAsync(ws, stream) != hipSuccess && rc == PYGW_OK) rc = PYGW_E_HIP;
  return rc;
}

static int minc_launch(bool second, const int* plan_host, const int* plan_dev, const double* kv, int bstart, int bend,
                       int cstart, int cend, const void* a, const void* b, const void* g, int nbmax, int nbmax_g,
                       int nlomax, const double* s3r, int n_mix, int n_fnc, const int* corind, int ncg_rows,
                       const double* pos, const double* cgcoef, int n_cg, void* out, long ld, int col0,
                       hipStream_t stream) {
  using namespace mm;
  const int nb = bend - bstart, nc = cend - cstart;
  if (!plan_ok(plan_host) || !plan_dev || !kv || !a || !b || !s3r || !corind || !pos || !cgcoef || !out || nb <= 0 ||
      nc <= 0 || bstart < 0 || bend > nbmax || cstart < 0 || cend > ncg_rows || n_mix <= 0 || n_fnc <= 0 ||
      n_cg <= 0 || col0 < 0 || ld < (long)col0 + plan_host[H_LOCT] || nlomax < plan_max_lo(plan_host) ||
      (nlomax > 0 && (!g || bend > nbmax_g)) || !plan_fits_s3r(plan_host, n_mix, n_fnc))
    return PYGW_E_ARG;
  const long n = (long)nb * nc * plan_host[H_LOCT];
  if (second)
    hipLaunchKernelGGL(pygw_minc_k<true>, dim3(ceil_div(n, 256)), dim3(256), 0, stream, plan_dev, kv, pos, bstart, nb,
                       cstart, nc, (const double2*)a, (const double2*)b, (const double2*)g, nbmax, nbmax_g, nlomax,
                       s3r, n_mix, n_fnc, corind, ncg_rows, cgcoef, n_cg, (double2*)out, ld, col0);
  else
    hipLaunchKernelGGL(pygw_minc_k<false>, dim3(ceil_div(n, 256)), dim3(256), 0, stream, plan_dev, kv, pos, bstart,
                       nb, cstart, nc, (const double2*)a, (const double2*)b, (const double2*)g, nbmax, nbmax_g, nlomax,
                       s3r, n_mix, n_fnc, corind, ncg_rows, cgcoef, n_cg, (double2*)out, ld, col0);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

extern "C" int pygw_minc(const int* plan_host, const int* plan_dev, const double* kl, int nstart, int nend,
                         int cstart, int cend, const void* alfa, const void* beta, const void* gama, int nbmax,
                         int nbmax_g, int nlomax, const double* s3r, int n_mix, int n_fnc, const int* corind,
                         int ncg_rows, const double* pos, const double* cgcoef, int n_cg, void* mmat, long ld_mmat,
                         int col0, void* stream) {
  return minc_launch(false, plan_host, plan_dev, kl, nstart, nend, cstart, cend, alfa, beta, gama, nbmax, nbmax_g,
                     nlomax, s3r, n_mix, n_fnc, corind, ncg_rows, pos, cgcoef, n_cg, mmat, ld_mmat, col0,
                     (hipStream_t)stream);
}

extern "C" int pygw_minc2(const int* plan_host, const int* plan_dev, const double* kmql, int cstart, int cend,
                          int mstart, int mend, const void* alfp, const void* betp, const void* gamp, int mbmax,
                          int mbmax_g, int nlomax, const double* s3r, int n_mix, int n_fnc, const int* corind,
                          int ncg_rows, const double* pos, const double* cgcoef, int n_cg, void* mmat, long ld_mmat,
                          int col0, void* stream) {
  return minc_launch(true, plan_host, plan_dev, kmql, mstart, mend, cstart, cend, alfp, betp, gamp, mbmax, mbmax_g,
                     nlomax, s3r, n_mix, n_fnc, corind, ncg_rows, pos, cgcoef, n_cg, mmat, ld_mmat, col0,
                     (hipStream_t)stream);
}
