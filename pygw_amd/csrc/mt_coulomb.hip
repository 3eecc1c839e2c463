// The muffin-tin blocks of the Coulomb matrix on gfx950 (include/pygw_hip.h):
//
//   pygw_ewald_summation: the structure constants of Ewald_Summation
//     (for_Coulomb.f90:1-164, called at pygw.py:2350) for every atom pair
//     i <= j,
//       sgm[lm][i + j(j+1)/2] = (S_R + S_G) / Gamma(l + 1/2)  (- 1/(eta pi), l = 0, i = j)
//       S_R = sum_{1e-6 < |R + r_ij| <= r_cf} Gamma_l(|R + r_ij|) e^{-i q.R'} Y_lm(R')
//       S_G = sum_{|q + G| <= g_cf} (4 pi^(3/2)/Vol) e^{-eta^2 |q+G|^2 / 4}
//             e^{i G.r_ij} |q+G|^-2 (-i |q+G|/2)^l Y_lm(q + G)
//     with R' = R + r_ij, Gamma_0 = sqrt(pi) erfc(r/eta)/r and
//     Gamma_l = Gamma_(l-1) (l - 1/2)/r + r^(l-2) e^{-r^2/eta^2} / eta^(2l-1).
//   pygw_mt_coulomb: MT_Coulomb (for_Coulomb.f90:166-283, pygw.py:2394), the
//     rows of the product functions of every atom (or of one atom type):
//       V[im][jm] = rtl_jrm rtl_irm sum_{m3,m4} (+-1) sgm(l1+l2, m3+m4)
//                   gtilde(l1, l2, m3, m4) conj(D^l1_{-m3,m1}(i)) D^l2_{m4,m2}(j)
//                   + delta 4 pi rrint(irm, jrm) / (2 l1 + 1)
//     (the sign and the conjugated sgm for i > j as in the reference).
//   pygw_radial_olap: the radial overlaps of the product functions with
//     j_L(k r) r (Calc_Olap_PW_Product, pygw.py:2109-2124: jlam) or with
//     sin(k r) (Zero_Momentum_Coulomb, pygw.py:2081-2086), integrated by the
//     rule of RINT13g (radials.f90:140-187) as a weighted sum.
//   pygw_zero_momentum_coulomb: the q = 0 term of the L = 0 functions
//     (pygw.py:2055-2104): phase[i][j][g] = sum over G of length g of
//     e^{-2 pi i G.(pos_i - pos_j)} (pw_get_short_and_phase,
//     for_Coulomb.f90:515-582), then
//       V[im][jm] = sum_g phase[i][j][g] (16 pi^2/Vol) / G_g^4 s_im(g) s_jm(g).
//
// Sums run in a fixed order per thread and are reduced in a fixed order, so
// every result is the same from run to run.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "../../include/pygw_hip.h"

namespace mt {
constexpr int kMaxLam = PYGW_EWALD_MAX_LAM;  // l of the structure constants
constexpr double kMaxCells = 64.0;           // cutoff / shortest basis vector
constexpr double kPi = 3.14159265358979323846;
// the value of pi that the reference's Ewald_Summation, MT_Coulomb and
// pw_get_short_and_phase carry (for_Coulomb.f90:25, 196, 532)
constexpr double kPiF = 3.14159265358979;
}  // namespace mt

// The Ewald sums run over vector lists that the host builds as genr_str
// does (for_Coulomb.f90:418-491): candidates (i1, i2, i3) in [-imax, imax]^3,
// imax = int(cutoff / shortest basis vector) + 1, kept for 1e-6 < |v| <=
// cutoff, ordered by length (ties by generation order).
//   rvec[4 (roff[pair] + n)] = (R + r_ij, |R + r_ij|), the real-space list of
//     each atom pair;
//   gvec[8 n] = (q + G, |q + G|, G), the reciprocal list (the same for every
//     pair).
//
// tab[n][lm] = c_l(v_n) Y_lm(v_n) for every vector of a list, one thread per
// (vector, m >= 0) that runs l = m .. lam and writes (l, m) and (l, -m):
//   kind 0 (real space, v = R + r_ij, r = |v|): c_l = Gamma_l(r) e^{-i q.v},
//     Gamma_0 = sqrt(pi) erfc(r/eta)/r,
//     Gamma_l = Gamma_(l-1) (l - 1/2)/r + r^(l-2) e^{-r^2/eta^2} / eta^(2l-1);
//   kind 1 (reciprocal space, v = q + G, g = |v|):
//     c_l = pref e^{-eta^2 g^2/4} g^-2 (g/2)^l (-i)^l.
// Y_lm by the normalised associated Legendre recurrences
//   P_mm = -sqrt((2m+1)/(2m)) sin(t) P_(m-1)(m-1),  P_(m+1)m = sqrt(2m+3) cos(t) P_mm,
//   P_lm = a_lm (cos(t) P_(l-1)m - b_lm P_(l-2)m),
// times e^{i m phi}, Y_l,-m = (-1)^m conj(Y_lm); v on the z axis takes phi = 0.
__global__ void __launch_bounds__(256) pygw_ew_tab(int kind, const double* __restrict__ vec, int vstride, int nvec,
                                                   int lam, const double* __restrict__ pq, double eta, double pref,
                                                   double2* __restrict__ tab) {
  using namespace mt;
  const long e = blockIdx.x * 256L + threadIdx.x;
  const int m = (int)(e % (lam + 1));
  const long n = e / (lam + 1);
  if (n >= nvec) return;
  const int nlm = (lam + 1) * (lam + 1);
  const double* v = vec + (size_t)vstride * n;
  const double r = v[3];
  const double rxy = sqrt(v[0] * v[0] + v[1] * v[1]);
  const double ct = v[2] / r, st = rxy / r;
  const double cp = rxy > 0.0 ? v[0] / rxy : 1.0, sp = rxy > 0.0 ? v[1] / rxy : 0.0;
  // P_mm and e^{i m phi}
  double pmm = 1.0 / sqrt(4.0 * kPi), cm = 1.0, sm = 0.0;
  for (int k = 1; k <= m; k++) {
    pmm *= -sqrt((2.0 * k + 1.0) / (2.0 * k)) * st;
    const double c2 = cm * cp - sm * sp, s2 = sm * cp + cm * sp;
    cm = c2;
    sm = s2;
  }
  // c_l at l = 0, and its step to l + 1
  double2 ph;        // the l-independent complex factor
  double cl, gam = 0.0, rpow = 0.0, epow = 0.0, gausr = 0.0;
  if (kind == 0) {
    double sq, cq;
    sincos(-(v[0] * pq[0] + v[1] * pq[1] + v[2] * pq[2]), &sq, &cq);
    ph = double2{cq, sq};
    gausr = exp(-r * r / (eta * eta));
    gam = sqrt(kPiF) * erfc(r / eta) / r;
    rpow = 1.0 / r;
    epow = eta;
    for (int ll = 1; ll <= m; ll++) {
      gam = gam * (ll - 0.5) / r + rpow * gausr / epow;
      rpow *= r;
      epow *= eta * eta;
    }
    cl = gam;
  } else {
    ph = double2{1.0, 0.0};
    cl = pref * exp(-(eta * r) * (eta * r) / 4.0) / (r * r);
    for (int k = 0; k < m; k++) cl *= 0.5 * r;
  }
  double2* out = tab + (size_t)n * nlm;
  const double sg = (m & 1) ? -1.0 : 1.0;
  double p2 = 0.0, p1 = pmm;
  for (int l = m; l <= lam; l++) {
    double p;
    if (l == m) {
      p = pmm;
    } else {
      if (l == m + 1) {
        p = sqrt(2.0 * m + 3.0) * ct * pmm;
      } else {
        const double aa = sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
        const double bb = sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
        p = aa * (ct * p1 - bb * p2);
      }
      p2 = p1;
      p1 = p;
      if (kind == 0) {
        gam = gam * (l - 0.5) / r + rpow * gausr / epow;
        rpow *= r;
        epow *= eta * eta;
        cl = gam;
      } else {
        cl *= 0.5 * r;
      }
    }
    // z = c_l ph Y_lm (Y_lm = p e^{i m phi}), and for -m: c_l ph (-1)^m conj(p e^{i m phi})
    double2 f{cl * ph.x, cl * ph.y};
    if (kind == 1) {  // times (-i)^l
      switch (l & 3) {
        case 1: f = double2{f.y, -f.x}; break;
        case 2: f = double2{-f.x, -f.y}; break;
        case 3: f = double2{-f.y, f.x}; break;
        default: break;
      }
    }
    const double yr = p * cm, yi = p * sm;
    out[l * l + l + m] = double2{f.x * yr - f.y * yi, f.x * yi + f.y * yr};
    if (m > 0) out[l * l + l - m] = double2{sg * (f.x * yr + f.y * yi), sg * (f.y * yr - f.x * yi)};
  }
}

// eph[pair][n] = e^{i G_n.r_ij}: one thread per entry.
__global__ void __launch_bounds__(256) pygw_ew_phase(const double* __restrict__ gvec, int ng,
                                                     const double* __restrict__ raa, int npair,
                                                     double2* __restrict__ eph) {
  const long e = blockIdx.x * 256L + threadIdx.x;
  if (e >= (long)ng * npair) return;
  const int n = (int)(e % ng), pair = (int)(e / ng);
  const double* v = gvec + 8 * (size_t)n;
  double sn, cn;
  sincos(v[4] * raa[3 * pair] + v[5] * raa[3 * pair + 1] + v[6] * raa[3 * pair + 2], &sn, &cn);
  eph[e] = double2{cn, sn};
}

// part[(c * npair + pair) * nlm + lm] = sum over the reciprocal vectors n of
// chunk c (in list order) of gtab[n][lm] eph[pair][n]: one thread per lm of a
// block of kEwPairs pairs and a chunk (blockIdx.z); each table entry is read
// once for the block's pairs, the phases are the same across the block.
constexpr int kEwPairs = 8;
__global__ void __launch_bounds__(64) pygw_ew_recip(const double2* __restrict__ gtab, const double2* __restrict__ eph,
                                                    int ng, int nlm, int npair, int nchunk,
                                                    double2* __restrict__ part) {
  const int lm = blockIdx.x * 64 + threadIdx.x, c = blockIdx.z, p0 = blockIdx.y * kEwPairs;
  const int np = min(kEwPairs, npair - p0);
  const int n0 = (int)((long)ng * c / nchunk), n1 = (int)((long)ng * (c + 1) / nchunk);
  if (lm >= nlm) return;
  double2 s[kEwPairs];
#pragma unroll
  for (int q = 0; q < kEwPairs; q++) s[q] = double2{0.0, 0.0};
  for (int n = n0; n < n1; n++) {
    const double2 t = gtab[(size_t)n * nlm + lm];
#pragma unroll
    for (int q = 0; q < kEwPairs; q++) {
      if (q < np) {
        const double2 e = eph[(size_t)(p0 + q) * ng + n];
        s[q].x += t.x * e.x - t.y * e.y;
        s[q].y += t.x * e.y + t.y * e.x;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < kEwPairs; q++)
    if (q < np) part[((size_t)c * npair + p0 + q) * nlm + lm] = s[q];
}

// sgm[lm + nlm pair] = (sum over the pair's real-space vectors of rtab[n][lm]
// + sum_c part[c]) / Gamma(l + 1/2), minus 1/(eta pi) for l = 0, i = j: one
// thread per (lm, pair), every sum in list / chunk order.
__global__ void __launch_bounds__(64) pygw_ew_final(const double2* __restrict__ rtab, const int* __restrict__ roff,
                                                    const double2* __restrict__ part, int nchunk, int npair, int lam,
                                                    double eta, double2* __restrict__ sgm) {
  using namespace mt;
  const int nlm = (lam + 1) * (lam + 1);
  const int pair = blockIdx.x, lm = blockIdx.y * 64 + threadIdx.x;
  if (lm >= nlm) return;
  int l = 0;
  while ((l + 1) * (l + 1) <= lm) l++;
  double2 s1{0.0, 0.0};
  for (int n = roff[pair]; n < roff[pair + 1]; n++) {
    const double2 t = rtab[(size_t)n * nlm + lm];
    s1.x += t.x;
    s1.y += t.y;
  }
  double2 s2{0.0, 0.0};
  for (int c = 0; c < nchunk; c++) {
    const double2 t = part[((size_t)c * npair + pair) * nlm + lm];
    s2.x += t.x;
    s2.y += t.y;
  }
  double gam = sqrt(kPiF);
  for (int k = 1; k <= l; k++) gam *= (k - 0.5);
  double2 v{(s1.x + s2.x) / gam, (s1.y + s2.y) / gam};
  int j = 0;
  while ((j + 1) * (j + 2) / 2 <= pair) j++;
  if (lm == 0 && pair - j * (j + 1) / 2 == j) v.x -= 1.0 / (eta * kPiF);
  sgm[lm + (size_t)nlm * pair] = v;
}

// ---------------------------------------------------------------------------
// MT_Coulomb.  Product function im (global row index) -> (atom idf, type
// iat, function irm, l, m): the functions of type a start at im_start[a],
// its mult[a] equivalent atoms follow one another, each with the functions
// irm = 0 .. nmix[a]-1 and m = -l .. l.
struct MtRow {
  int idf, iat, irm, l, m;
};

__device__ inline MtRow row_of(int im, int nat, int max_nmix, const int* __restrict__ big_l, const int* __restrict__ nmix,
                               const int* __restrict__ im_start, const int* __restrict__ mult) {
  int a = 0, idf0 = 0;
  while (a + 1 < nat && im >= im_start[a + 1]) {
    idf0 += mult[a];
    a++;
  }
  int per = 0;
  for (int f = 0; f < nmix[a]; f++) per += 2 * big_l[f + max_nmix * a] + 1;
  const int off = im - im_start[a];
  int r = off % per, f = 0;
  while (f + 1 < nmix[a] && r >= 2 * big_l[f + max_nmix * a] + 1) {
    r -= 2 * big_l[f + max_nmix * a] + 1;
    f++;
  }
  const int l = big_l[f + max_nmix * a];
  return MtRow{idf0 + off / per, a, f, l, r - l};
}

__device__ inline double gettildeg(int l1, int l2, int m1, int m2, const double* __restrict__ tilg, int tsize) {
  const int par = abs(l1 - l2) % 2;
  int j1, j2, mj1, mj2, fact;
  if (l1 < l2) {
    j1 = l2; j2 = l1; mj1 = m2; mj2 = m1; fact = 1 - 2 * par;
  } else {
    j1 = l1; j2 = l2; mj1 = m1; mj2 = m2; fact = 1;
  }
  if (mj2 < 0) {
    mj1 = -mj1;
    mj2 = -mj2;
  }
  const int i1 = j2 + mj2 + 1, i2 = (j2 + 1) * (mj1 + j1), i3 = (j2 + 1) * j1 * j2 + (j2 - 1) * j2 / 2;
  const int i4 = (j1 + 1) * (j1 + 2) * (3 * j1 - 1) * j1 / 12;
  const int i = i1 + i2 + i3 + i4 - 1;
  return i < tsize ? fact * tilg[i] : __builtin_nan("");
}

__device__ inline double2 getdjmm(int idf, int l, int m1, int m2, const double2* __restrict__ djmm, int dimdj) {
  const int sg = m2 >= 0 ? 1 : -1;
  const int i = l * (l + 1) * (4 * l - 1) / 6 + (l + 1) * (sg * m1 + l) + abs(m2);
  if (i >= dimdj) return double2{__builtin_nan(""), 0.0};
  const double2 d = djmm[i + (size_t)dimdj * idf];
  if (m2 >= 0) return d;
  const double s = (abs(m1 - m2) & 1) ? -1.0 : 1.0;
  return double2{s * d.x, -s * d.y};
}

__device__ inline double2 cmul(double2 a, double2 b) { return double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

// V[im + loct*jm] (Fortran) for the rows in [r0, r1); V is zero elsewhere
// (pygw_zero_v2 first).  One thread per entry.
__global__ void __launch_bounds__(256) pygw_mt_fill(const int* __restrict__ big_l, const int* __restrict__ nmix,
                                                    const int* __restrict__ im_start, const int* __restrict__ mult,
                                                    int nat, int max_nmix, int loct, int r0, int r1,
                                                    const double* __restrict__ vq,
                                                    const double* __restrict__ tilg, int tsize,
                                                    const double* __restrict__ rrint,
                                                    const double2* __restrict__ djmm, int dimdj,
                                                    const double* __restrict__ rtlij,
                                                    const double2* __restrict__ sgm, int lam,
                                                    double2* __restrict__ V) {
  const int nlm = (lam + 1) * (lam + 1);
  const long e = blockIdx.x * 256L + threadIdx.x;
  const long nr = (long)(r1 - r0) * loct;
  if (e >= nr) return;
  const int im = r0 + (int)(e % (r1 - r0)), jm = (int)(e / (r1 - r0));
  const bool q0 = (fabs(vq[0]) + fabs(vq[1]) + fabs(vq[2])) < 1e-7;
  const MtRow a = row_of(im, nat, max_nmix, big_l, nmix, im_start, mult);
  const MtRow b = row_of(jm, nat, max_nmix, big_l, nmix, im_start, mult);
  double2 out{0.0, 0.0};
  if (a.l + b.l > lam) {  // structure constants of too low an order
    out = double2{__builtin_nan(""), __builtin_nan("")};
  } else if (!(q0 && a.l == 0 && b.l == 0)) {
    const int ii = min(a.idf, b.idf), jj = max(a.idf, b.idf);
    const double2* sg = sgm + (size_t)nlm * (ii + jj * (jj + 1) / 2);
    const int L = a.l + b.l;
    double2 s{0.0, 0.0};
    for (int m3 = -a.l; m3 <= a.l; m3++) {
      const double2 d3 = getdjmm(a.idf, a.l, -m3, a.m, djmm, dimdj);
      const double2 d3c{d3.x, -d3.y};
      for (int m4 = -b.l; m4 <= b.l; m4++) {
        double2 stc;
        double mn;
        if (a.idf <= b.idf) {
          stc = sg[L * L + L + m3 + m4];
          mn = (abs(m3) & 1) ? -1.0 : 1.0;
        } else {
          const double2 t = sg[L * L + L - m3 - m4];
          stc = double2{t.x, -t.y};
          mn = (abs(m4 + L) & 1) ? -1.0 : 1.0;
        }
        const double tg = gettildeg(a.l, b.l, m3, m4, tilg, tsize);
        const double2 d4 = getdjmm(b.idf, b.l, m4, b.m, djmm, dimdj);
        const double2 p = cmul(cmul(stc, d3c), d4);
        s.x += mn * tg * p.x;
        s.y += mn * tg * p.y;
      }
    }
    // rtlij(jrm, irm, jat, iat), Fortran [max_nmix][max_nmix][nat][nat]
    const double rt = rtlij[b.irm + (size_t)max_nmix * (a.irm + (size_t)max_nmix * (b.iat + (size_t)nat * a.iat))];
    out = double2{rt * s.x, rt * s.y};
    if (a.m == b.m && a.l == b.l && a.idf == b.idf) {
      const int i1 = min(a.irm, b.irm), j1 = max(a.irm, b.irm);
      out.x += 4.0 * mt::kPiF * rrint[a.iat + (size_t)nat * (j1 * (j1 + 1) / 2 + i1)] / (2.0 * a.l + 1.0);
    }
  }
  V[im + (size_t)loct * jm] = out;
}

// ---------------------------------------------------------------------------
// Radial overlaps.  out[f][k] = sum_j w_j (A_f(j) + c2 B_f(j)) F(k r_j), F =
// j_L(k r) r (mode 0, L = lf[f]) or sin(k r) (mode 1) on the caller's mesh r_j
// (rx), with the RINT13g weights w_j of the logarithmic mesh (pygw_rad_olap).
//
// j_L: the power series x^L/(2L+1)!! (1 - y/(1!(2L+3)) + y^2/(2!(2L+3)(2L+5)) - ...),
// y = x^2/2, below x = 0.01 (the first omitted term is below 1e-20
// relative); for x > L the upward recurrence j_(l+1) = (2l+1)/x j_l - j_(l-1)
// from j_0 = sin x / x and j_1 = sin x / x^2 - cos x / x (stable while l < x);
// otherwise Miller's downward recurrence j_(l-1) = (2l+1)/x j_l - j_(l+1)
// started at order L + 40 + x (<= 2L + 40) from (0, 1e-100) (values only grow
// from there while l > x, and their squares stay normal), normalised with
// the identity sum_l (2l+1) j_l^2 = 1, its sign fixed by comparing j_0 =
// sin x / x or j_-1 = cos x / x (whichever is larger) with the recurrence's
// own values of those two.
__device__ double sph_bessel(int L, double x) {
  if (x < 0.01) {
    const double y = 0.5 * x * x;
    double df = 1.0;  // (2L+1)!!
    for (int k = 3; k <= 2 * L + 1; k += 2) df *= k;
    double term = 1.0, sum = 1.0;
    for (int k = 1; k <= 4; k++) {
      term *= -y / (k * (2.0 * L + 2.0 * k + 1.0));
      sum += term;
    }
    double xl = 1.0;
    for (int k = 0; k < L; k++) xl *= x;
    return xl / df * sum;
  }
  double sx, cx;
  sincos(x, &sx, &cx);
  if (x > (double)L) {
    double jm = sx / x;  // j_0
    if (L == 0) return jm;
    double jc = (sx / x - cx) / x;  // j_1
    for (int l = 1; l < L; l++) {
      const double jn = (2.0 * l + 1.0) / x * jc - jm;
      jm = jc;
      jc = jn;
    }
    return jc;
  }
  const int N = L + 40 + (int)x;
  const double ix = 1.0 / x;
  double jp = 0.0, jc = 1e-100, norm = 0.0, want = 0.0;
  for (int l = N; l >= 0; l--) {
    if (l == L) want = jc;
    norm += (2.0 * l + 1.0) * jc * jc;
    const double jm = (2.0 * l + 1.0) * ix * jc - jp;
    jp = jc;
    jc = jm;
    if (fabs(jc) > 1e150) {  // rescale every kept value alike
      jc *= 1e-150;
      jp *= 1e-150;
      want *= 1e-150;
      norm *= 1e-300;
    }
  }
  // jp = j_0, jc = j_-1 (both unnormalised)
  const double sgn = fabs(sx) >= fabs(cx) ? (jp * sx < 0.0 ? -1.0 : 1.0) : (jc * cx < 0.0 ? -1.0 : 1.0);
  return sgn * want / sqrt(norm);
}


// w[j] = RINT13g's rule (radials.f90:140-187) for npt >= 4 points as the
// weight of point j: Simpson's weights dh r_j (1, 4, 2, ..., 4, 1)/3 from the
// second point (1-based, odd npt) or the third (even npt, with a trapezoid
// dh (f_1 + f_2)/2 over the first interval), and r0 f(r0)/3 for [0, r0].  The
// r_j are the ones RINT13g forms: r0 D^(J0-1) at the first Simpson point J0
// and a product by D per point after it (which drifts from r0 D^j by up to
// ~1e-13 at the end of an 800-point mesh), r0 D^(J0-1) / D at point J0 - 1.
// One thread walks the mesh in order (thread 0 of pygw_rad_table's first
// block, alongside the table).
__device__ void rad_weights(int npt, const double* __restrict__ rx, double dh, double D, double* __restrict__ w) {
  const double r0 = rx[0];
  const bool odd = (npt & 1) != 0;
  const int j0 = odd ? 1 : 2;                // 0-based first Simpson point
  double R = odd ? r0 * D : r0 * (D * D);     // its r
  const double R1 = R / D;                    // r of point j0 - 1
  const double dh3 = dh / 3.0;
  w[0] = r0 / 3.0 + (odd ? dh3 * R1 : 0.5 * dh * r0);
  if (!odd) w[1] = dh3 * R1 + 0.5 * dh * R1;
  // pairs of points (4, 2) from j0 on; the last point has weight 1
  const double c4 = 4.0 * dh3, c2 = 2.0 * dh3;
  int j = j0;
  for (; j + 2 < npt; j += 2) {
    w[j] = c4 * R;
    R *= D;
    w[j + 1] = c2 * R;
    R *= D;
  }
  for (; j < npt; j++) {
    w[j] = (j == npt - 1 ? dh3 : ((j - j0) & 1) == 0 ? c4 : c2) * R;
    R *= D;
  }
}

// tab[(L*npt + j)*nk + k] = j_L(k r_j) r_j (mode 0, L < nl) or sin(k r_j)
// (mode 1, nl = 1): one thread per (j, k), consecutive threads along k; the
// orders below x by one upward recurrence from j_0, j_1, the others by
// sph_bessel.
__global__ void __launch_bounds__(256) pygw_rad_table(int mode, int nl, int npt, const double* __restrict__ rx,
                                                      const double* __restrict__ k, int nk, double* __restrict__ tab,
                                                      double dh, double D, double* __restrict__ w) {
  const long e = blockIdx.x * 256L + threadIdx.x;
  if (e == 0) rad_weights(npt, rx, dh, D, w);
  if (e >= (long)npt * nk) return;
  const int ik = (int)(e % nk), j = (int)(e / nk);
  const double r = rx[j], x = k[ik] * r;
  const size_t stride = (size_t)npt * nk;
  if (mode == 1) {
    tab[e] = sin(x);
    return;
  }
  if (x < 0.01) {
    for (int L = 0; L < nl; L++) tab[e + L * stride] = sph_bessel(L, x) * r;
    return;
  }
  double sx, cx;
  sincos(x, &sx, &cx);
  double ja = sx / x, jb = (sx / x - cx) / x;  // j_L, j_(L+1) at L = 0
  for (int L = 0; L < nl; L++) {
    tab[e + L * stride] = ((double)L < x ? ja : sph_bessel(L, x)) * r;
    if ((double)(L + 1) < x) {  // j_(L+2) while it stays below x
      const double jn = (2.0 * L + 3.0) / x * jb - ja;
      ja = jb;
      jb = jn;
    }
  }
}

// out[f][k] = sum_j w_j (A_f(j) + c2 B_f(j)) tab[L_f][j][k]: a block of
// kOlapWaves waves covers 64 k (the lanes) and kFun consecutive functions
// (blockIdx.y), wave v the points j in [v npt / kOlapWaves, (v+1) npt /
// kOlapWaves).  The block first stages W[q][j] = w_j (A + c2 B) of its
// functions in LDS (npt <= kOlapMaxPts); the waves' partial sums are added in
// wave order through LDS.  A table row is read once for consecutive
// functions of the same L.  L_f outside [0, nl) gives NaN.
constexpr int kFun = 4, kOlapWaves = 16, kOlapMaxPts = 1024;
__global__ void __launch_bounds__(64 * kOlapWaves) pygw_rad_olap(const double* __restrict__ ul,
                                                                 const double* __restrict__ us,
                                                                 const int* __restrict__ lf, int nf, int nl, int npt,
                                                                 const double* __restrict__ w, double c2,
                                                                 const double* __restrict__ tab, int nk,
                                                                 double* __restrict__ out) {
  __shared__ double W[kFun][kOlapMaxPts];
  __shared__ double red[kOlapWaves - 1][kFun][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int f0 = blockIdx.y * kFun, ik = blockIdx.x * 64 + lane;
  const int nq = min(kFun, nf - f0);
  int L[kFun];
  bool bad = false;
#pragma unroll
  for (int q = 0; q < kFun; q++) {
    L[q] = q < nq ? lf[f0 + q] : -1;
    if (q < nq && (L[q] < 0 || L[q] >= nl)) bad = true;
  }
  for (int e = threadIdx.x; e < nq * npt; e += 64 * kOlapWaves) {
    const int q = e / npt, j = e % npt;
    const size_t o = (size_t)(f0 + q) * npt + j;
    W[q][j] = w[j] * (ul[o] + c2 * us[o]);
  }
  __syncthreads();
  double acc[kFun];
#pragma unroll
  for (int q = 0; q < kFun; q++) acc[q] = 0.0;
  const size_t stride = (size_t)npt * nk;
  const int j0 = (int)((long)npt * wv / kOlapWaves), j1 = (int)((long)npt * (wv + 1) / kOlapWaves);
  if (!bad && ik < nk) {
    for (int j = j0; j < j1; j++) {
      const double* t = tab + (size_t)j * nk + ik;
      double tv = 0.0;
      int cur = -1;
#pragma unroll
      for (int q = 0; q < kFun; q++) {
        if (q < nq) {
          if (L[q] != cur) {
            cur = L[q];
            tv = t[cur * stride];
          }
          acc[q] += W[q][j] * tv;
        }
      }
    }
  }
  if (wv > 0) {
#pragma unroll
    for (int q = 0; q < kFun; q++) red[wv - 1][q][lane] = acc[q];
  }
  __syncthreads();
  if (wv == 0 && ik < nk) {
    for (int v = 1; v < kOlapWaves; v++) {
#pragma unroll
      for (int q = 0; q < kFun; q++) acc[q] += red[v - 1][q][lane];
    }
#pragma unroll
    for (int q = 0; q < kFun; q++)
      if (q < nq) out[(size_t)(f0 + q) * nk + ik] = bad ? __builtin_nan("") : acc[q];
  }
}

// phase[(i*ndf + j)*ngs + g] = sum over the G of group g (gind4, 1-based, G 0
// excluded: ipw >= 1) of e^{-2 pi i G.(pos_i - pos_j)} (1 for i = j).  Groups
// are runs of equal length in the length-sorted G list: g0[g] .. g0[g+1].
__global__ void __launch_bounds__(256) pygw_zm_phase(const int* __restrict__ gindex, int ng, const int* __restrict__ g0,
                                                     int ngs, const double* __restrict__ pos, int ndf,
                                                     double2* __restrict__ phase) {
  const long e = blockIdx.x * 256L + threadIdx.x;
  if (e >= (long)ndf * ndf * ngs) return;
  const int g = (int)(e % ngs), ij = (int)(e / ngs), i = ij / ndf, j = ij % ndf;
  double2 s{0.0, 0.0};
  const double d0 = pos[0 + 3 * i] - pos[0 + 3 * j], d1 = pos[1 + 3 * i] - pos[1 + 3 * j], d2 = pos[2 + 3 * i] - pos[2 + 3 * j];
  const int p0 = max(g0[g], 0), p1 = min(g0[g + 1], ng);
  for (int p = p0; p < p1; p++) {
    if (i == j) {
      s.x += 1.0;
      continue;
    }
    const double a = -2.0 * mt::kPiF * (gindex[p] * d0 + gindex[p + ng] * d1 + gindex[p + 2 * ng] * d2);
    double sa, ca;
    sincos(a, &sa, &ca);
    s.x += ca;
    s.y += sa;
  }
  phase[e] = s;
}

// V[im + mbs*jm] for the L = 0 functions: sum_g phase[i][j][g] c/G^4 s_a(g) s_b(g).
// One wave per entry (four a block): lane l sums g = l, l + 64, ..., then the
// lanes are added by a fixed butterfly, so the result does not depend on
// timing.
__global__ void __launch_bounds__(256) pygw_zm_fill(const int* __restrict__ fidx, const int* __restrict__ fdf,
                                                    const int* __restrict__ frow, int nf0, const double* __restrict__ sing,
                                                    int nsing,
                                                    const double* __restrict__ glen0, int ngs,
                                                    const double2* __restrict__ phase, int ndf, double vol, int mbs,
                                                    double2* __restrict__ V) {
  const int e = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (e >= nf0 * nf0) return;
  const int a = e % nf0, b = e / nf0;
  if (frow[a] < 0 || frow[a] >= mbs || frow[b] < 0 || frow[b] >= mbs) return;
  if (fidx[a] < 0 || fidx[a] >= nsing || fidx[b] < 0 || fidx[b] >= nsing || fdf[a] < 0 || fdf[a] >= ndf ||
      fdf[b] < 0 || fdf[b] >= ndf) {
    if (lane == 0) V[frow[a] + (size_t)mbs * frow[b]] = double2{__builtin_nan(""), __builtin_nan("")};
    return;
  }
  const double c = 16.0 * mt::kPi * mt::kPi / vol;
  const double* sa = sing + (size_t)fidx[a] * ngs;
  const double* sb = sing + (size_t)fidx[b] * ngs;
  const double2* ph = phase + ((size_t)fdf[a] * ndf + fdf[b]) * ngs;
  double2 s{0.0, 0.0};
  for (int g = lane; g < ngs; g += 64) {
    const double gl2 = glen0[g] * glen0[g];
    const double w = c / (gl2 * gl2) * sa[g] * sb[g];
    s.x += ph[g].x * w;
    s.y += ph[g].y * w;
  }
  for (int o = 32; o > 0; o >>= 1) {
    s.x += __shfl_xor(s.x, o);
    s.y += __shfl_xor(s.y, o);
  }
  if (lane == 0) V[frow[a] + (size_t)mbs * frow[b]] = s;
}

__global__ void pygw_zero_v2(double2* __restrict__ x, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    x[e] = double2{0.0, 0.0};
}

// ---------------------------------------------------------------------------
namespace {
// genr_str (for_Coulomb.f90:418-491): basis[k] (k < 3) the basis vectors,
// the vectors i1 b0 + i2 b1 + i3 b2 + shift (i1 slowest) with 1e-6 < |v| <=
// rmax, as (x, y, z, |v|), ordered by length (ties in generation order).
// False if rmax spans more than kMaxCells shortest basis vectors.
bool lattice_list(double rmax, const double shift[3], const double basis[3][3], std::vector<double>& out) {
  double lmin = 1e300;
  for (int k = 0; k < 3; k++)
    lmin = std::fmin(lmin, std::sqrt(basis[k][0] * basis[k][0] + basis[k][1] * basis[k][1] + basis[k][2] * basis[k][2]));
  if (!(lmin > 0.0) || !(rmax / lmin < mt::kMaxCells)) return false;
  const int imax = (int)(rmax / lmin) + 1;
  std::vector<std::array<double, 4>> v;
  for (int i1 = -imax; i1 <= imax; i1++)
    for (int i2 = -imax; i2 <= imax; i2++)
      for (int i3 = -imax; i3 <= imax; i3++) {
        std::array<double, 4> r;
        for (int x = 0; x < 3; x++) r[x] = basis[0][x] * i1 + basis[1][x] * i2 + basis[2][x] * i3 + shift[x];
        r[3] = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        if (r[3] <= rmax && r[3] > 1e-6) v.push_back(r);
      }
  std::stable_sort(v.begin(), v.end(), [](const std::array<double, 4>& a, const std::array<double, 4>& b) {
    return a[3] < b[3];
  });
  out.clear();
  for (const auto& r : v) out.insert(out.end(), r.begin(), r.end());
  return true;
}

struct EwaldPlan {
  std::vector<double> rvec, gvec, raa;  // the lists (4 / 8 doubles a vector) and r_ij per pair
  std::vector<int> roff;
  int ng = 0, npair = 0, nlm = 0, nchunk = 1;
  size_t off_gvec = 0, off_raa = 0, off_pq = 0, off_roff = 0, off_gtab = 0, off_rtab = 0, off_eph = 0, off_part = 0;
  size_t bytes = 0;
};

// The vector lists on the host and the workspace layout; false on a bad argument.
bool ewald_plan(const double* pq, int lam_max, const double* pos, const double* br2, const double* rbas,
                const double* alat, int ortho, double r_cf, double g_cf, int ndf, EwaldPlan& P) {
  if (!pq || !pos || !br2 || !rbas || !alat || lam_max < 0 || lam_max > mt::kMaxLam || ndf <= 0 || ndf > 1000 ||
      !(r_cf > 0.0) || !(g_cf > 0.0) || (ortho != 0 && ortho != 1))
    return false;
  P.npair = ndf * (ndf + 1) / 2;
  P.nlm = (lam_max + 1) * (lam_max + 1);
  double rb[3][3], gb[3][3];  // rows of rbas (Fortran rbas(k, x)); columns of br2 (br2(x, k))
  for (int k = 0; k < 3; k++)
    for (int x = 0; x < 3; x++) {
      rb[k][x] = rbas[k + 3 * x];
      gb[k][x] = br2[x + 3 * k];
    }
  P.rvec.clear();
  P.raa.assign(3 * (size_t)P.npair, 0.0);
  P.roff.assign(P.npair + 1, 0);
  std::vector<double> one;
  for (int j = 0; j < ndf; j++)
    for (int i = 0; i <= j; i++) {
      const int pair = i + j * (j + 1) / 2;
      double d[3], raa[3];
      for (int x = 0; x < 3; x++) d[x] = pos[x + 3 * i] - pos[x + 3 * j];
      for (int x = 0; x < 3; x++) raa[x] = ortho ? d[x] * alat[x] : d[0] * rb[0][x] + d[1] * rb[1][x] + d[2] * rb[2][x];
      for (int x = 0; x < 3; x++) P.raa[3 * pair + x] = raa[x];
      if (!lattice_list(r_cf, raa, rb, one)) return false;
      P.roff[pair] = (int)(P.rvec.size() / 4);  // pairs are visited in column order: pair = running index
      P.rvec.insert(P.rvec.end(), one.begin(), one.end());
    }
  P.roff[P.npair] = (int)(P.rvec.size() / 4);
  if (!lattice_list(g_cf, pq, gb, one)) return false;
  P.ng = (int)(one.size() / 4);
  P.gvec.assign(8 * (size_t)P.ng, 0.0);
  for (int n = 0; n < P.ng; n++) {
    for (int x = 0; x < 4; x++) P.gvec[8 * (size_t)n + x] = one[4 * (size_t)n + x];
    for (int x = 0; x < 3; x++) P.gvec[8 * (size_t)n + 4 + x] = one[4 * (size_t)n + x] - pq[x];
  }
  // chunks of the reciprocal sum: about 2048 blocks, at least 32 vectors a chunk
  const int tiles = ceil_div(P.npair, kEwPairs) * ceil_div(P.nlm, 64);
  P.nchunk = std::max(1, std::min({64, ceil_div(2048, tiles), P.ng / 32}));
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t o = al(P.rvec.size() * sizeof(double) + 8);
  P.off_gvec = o;
  o = al(o + P.gvec.size() * sizeof(double) + 8);
  P.off_raa = o;
  o = al(o + P.raa.size() * sizeof(double));
  P.off_pq = o;
  o = al(o + 3 * sizeof(double));
  P.off_roff = o;
  o = al(o + P.roff.size() * sizeof(int));
  P.off_gtab = o;
  o = al(o + (size_t)std::max(P.ng, 1) * P.nlm * sizeof(double2));
  P.off_rtab = o;
  o = al(o + (size_t)std::max((int)(P.rvec.size() / 4), 1) * P.nlm * sizeof(double2));
  P.off_eph = o;
  o = al(o + (size_t)std::max(P.ng, 1) * P.npair * sizeof(double2));
  P.off_part = o;
  o = al(o + (size_t)P.nchunk * P.npair * P.nlm * sizeof(double2));
  P.bytes = o;
  return true;
}
}  // namespace

extern "C" int pygw_ewald_workspace(const double* pq, int lam_max, const double* pos, const double* br2,
                                    const double* rbas, const double* alat, int ortho, double r_cf, double g_cf,
                                    int ndf, size_t* bytes) {
  EwaldPlan P;
  if (!bytes || !ewald_plan(pq, lam_max, pos, br2, rbas, alat, ortho, r_cf, g_cf, ndf, P)) return PYGW_E_ARG;
  *bytes = P.bytes;
  return PYGW_OK;
}

extern "C" int pygw_ewald_summation(const double* pq, int lam_max, const double* pos, const double* br2,
                                    const double* rbas, const double* alat, double vol, int ortho, double r_cf,
                                    double g_cf, double eta, int ndf, void* sgm, void* workspace, size_t ws_bytes,
                                    void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  EwaldPlan P;
  if (!sgm || !(vol > 0.0) || !(eta > 0.0) ||
      !ewald_plan(pq, lam_max, pos, br2, rbas, alat, ortho, r_cf, g_cf, ndf, P))
    return PYGW_E_ARG;
  char* ws = (char*)workspace;
  bool own = false;
  if (!ws) {
    HIP_TRY(hipMallocAsync((void**)&ws, P.bytes, stream));
    own = true;
  } else if (ws_bytes < P.bytes) {
    return PYGW_E_ARG;
  }
  // the lists travel in one host buffer; the stream is synchronised before it is freed
  std::vector<char> host(P.off_gtab, 0);
  std::memcpy(host.data(), P.rvec.data(), P.rvec.size() * sizeof(double));
  std::memcpy(host.data() + P.off_gvec, P.gvec.data(), P.gvec.size() * sizeof(double));
  std::memcpy(host.data() + P.off_raa, P.raa.data(), P.raa.size() * sizeof(double));
  std::memcpy(host.data() + P.off_pq, pq, 3 * sizeof(double));
  std::memcpy(host.data() + P.off_roff, P.roff.data(), P.roff.size() * sizeof(int));
  int rc = PYGW_OK;
  if (hipMemcpyAsync(ws, host.data(), host.size(), hipMemcpyHostToDevice, stream) != hipSuccess) rc = PYGW_E_HIP;
  const double* rvec = (const double*)ws;
  const double* gvec = (const double*)(ws + P.off_gvec);
  const double* raa = (const double*)(ws + P.off_raa);
  const double* pqd = (const double*)(ws + P.off_pq);
  const int* roff = (const int*)(ws + P.off_roff);
  double2* gtab = (double2*)(ws + P.off_gtab);
  double2* rtab = (double2*)(ws + P.off_rtab);
  double2* eph = (double2*)(ws + P.off_eph);
  double2* part = (double2*)(ws + P.off_part);
  const double pref = 4.0 * mt::kPiF * std::sqrt(mt::kPiF) / vol;
  const int nr = (int)(P.rvec.size() / 4);
  if (rc == PYGW_OK && P.ng > 0) {
    hipLaunchKernelGGL(pygw_ew_tab, dim3(ceil_div((long)P.ng * (lam_max + 1), 256)), dim3(256), 0, stream, 1, gvec, 8,
                       P.ng, lam_max, pqd, eta, pref, gtab);
    hipLaunchKernelGGL(pygw_ew_phase, dim3(ceil_div((long)P.ng * P.npair, 256)), dim3(256), 0, stream, gvec, P.ng, raa,
                       P.npair, eph);
  }
  if (rc == PYGW_OK && nr > 0)
    hipLaunchKernelGGL(pygw_ew_tab, dim3(ceil_div((long)nr * (lam_max + 1), 256)), dim3(256), 0, stream, 0, rvec, 4, nr,
                       lam_max, pqd, eta, pref, rtab);
  if (rc == PYGW_OK)
    hipLaunchKernelGGL(pygw_ew_recip, dim3(ceil_div(P.nlm, 64), ceil_div(P.npair, kEwPairs), P.nchunk), dim3(64), 0,
                       stream, gtab, eph, P.ng, P.nlm, P.npair, P.nchunk, part);
  if (rc == PYGW_OK)
    hipLaunchKernelGGL(pygw_ew_final, dim3(P.npair, ceil_div(P.nlm, 64)), dim3(64), 0, stream, rtab, roff, part,
                       P.nchunk, P.npair, lam_max, eta, (double2*)sgm);
  if (rc == PYGW_OK && hipGetLastError() != hipSuccess) rc = PYGW_E_HIP;
  if (own && hipFreeAsync(ws, stream) != hipSuccess && rc == PYGW_OK) rc = PYGW_E_HIP;
  if (hipStreamSynchronize(stream) != hipSuccess && rc == PYGW_OK) rc = PYGW_E_HIP;
  return rc;
}

extern "C" int pygw_mt_coulomb(const double* vq, int iat, int loctmatsize, const int* big_l, const int* nmix,
                               const int* im_start, const int* im_start_host, const int* mult, int nat, int max_nmix,
                               const double* tilg, int tsize, const double* rrint, const void* djmm, int dimdj,
                               const double* rtlij, const void* sgm, int lam_max, void* vmat, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (!vq || !big_l || !nmix || !im_start || !im_start_host || !mult || !tilg || !rrint || !djmm || !rtlij || !sgm ||
      !vmat || nat <= 0 || max_nmix <= 0 || loctmatsize <= 0 || iat < -1 || iat >= nat || dimdj <= 0 || tsize <= 0 ||
      lam_max < 0 || lam_max > mt::kMaxLam)
    return PYGW_E_ARG;
  const int r0 = iat < 0 ? 0 : im_start_host[iat], r1 = iat < 0 ? loctmatsize : im_start_host[iat + 1];
  if (r0 < 0 || r1 > loctmatsize || r1 < r0) return PYGW_E_ARG;
  if (iat >= 0)  // rows of the other atom types are zero
    hipLaunchKernelGGL(pygw_zero_v2, dim3(ceil_div((long)loctmatsize * loctmatsize, 256) < 4096
                                              ? ceil_div((long)loctmatsize * loctmatsize, 256) : 4096),
                       dim3(256), 0, stream, (double2*)vmat, (long)loctmatsize * loctmatsize);
  const long n = (long)(r1 - r0) * loctmatsize;
  if (n > 0)
    hipLaunchKernelGGL(pygw_mt_fill, dim3(ceil_div(n, 256)), dim3(256), 0, stream, big_l, nmix, im_start, mult, nat,
                       max_nmix, loctmatsize, r0, r1, vq, tilg, tsize, rrint, (const double2*)djmm, dimdj, rtlij,
                       (const double2*)sgm, lam_max, (double2*)vmat);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

extern "C" int pygw_radial_olap_workspace(int mode, int lmax, int npt, int n_k, size_t* bytes) {
  if ((mode != 0 && mode != 1) || lmax < 0 || (mode == 1 && lmax != 0) || npt < 4 || n_k <= 0 || !bytes)
    return PYGW_E_ARG;
  // the table of F, then the npt weights
  *bytes = ((size_t)(lmax + 1) * npt * n_k * sizeof(double) + 255) / 256 * 256 + (size_t)npt * sizeof(double);
  return PYGW_OK;
}

extern "C" int pygw_radial_olap(int mode, int rel, const double* ul, const double* us, const int* lf, int n_fun,
                                int lmax, int npt, const double* rx, double dh, const double* k, int n_k, double* out,
                                void* workspace, size_t ws_bytes, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  size_t need = 0;
  if (pygw_radial_olap_workspace(mode, lmax, npt, n_k, &need) != PYGW_OK || !ul || !us || !lf || !rx || !k || !out ||
      n_fun <= 0 || n_fun > 65535 || npt > kOlapMaxPts || !(dh > 0.0) || (rel != 0 && rel != 1))
    return PYGW_E_ARG;
  void* ws = workspace;
  bool own = false;
  if (!ws) {
    HIP_TRY(hipMallocAsync(&ws, need, stream));
    own = true;
  } else if (ws_bytes < need) {
    return PYGW_E_ARG;
  }
  // c2 = 1/c^2 (scalar relativistic) or RINT13g's single-precision literal 1e-22
  const double c2 = rel ? 1.0 / (137.0359895 * 137.0359895) : (double)1e-22f;
  // RINT13g's ratio e^dh from the host's libm, as the compiled reference
  // forms it (a last-bit difference in it grows j-fold along the mesh)
  const double D = std::exp(dh);
  const size_t tab_bytes = ((size_t)(lmax + 1) * npt * n_k * sizeof(double) + 255) / 256 * 256;
  double* wts = (double*)((char*)ws + tab_bytes);
  hipLaunchKernelGGL(pygw_rad_table, dim3(ceil_div((long)npt * n_k, 256)), dim3(256), 0, stream, mode, lmax + 1, npt,
                     rx, k, n_k, (double*)ws, dh, D, wts);
  hipLaunchKernelGGL(pygw_rad_olap, dim3(ceil_div(n_k, 64), ceil_div(n_fun, kFun)), dim3(64 * kOlapWaves), 0, stream,
                     ul, us, lf, n_fun, lmax + 1, npt, (const double*)wts, c2, (const double*)ws, n_k, out);
  int rc = hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
  if (own && hipFreeAsync(ws, stream) != hipSuccess && rc == PYGW_OK) rc = PYGW_E_HIP;
  return rc;
}

extern "C" int pygw_zero_momentum_workspace(int ndf, int n_short, size_t* bytes) {
  if (ndf <= 0 || n_short <= 0 || !bytes) return PYGW_E_ARG;
  *bytes = (size_t)ndf * ndf * n_short * sizeof(double2);
  return PYGW_OK;
}

extern "C" int pygw_zero_momentum_coulomb(const int* gindex, int n_g, const int* g_start, const double* glen0,
                                          int n_short, const double* pos, int ndf, const double* sing, int n_sing,
                                          const int* f_sing, const int* f_atom, const int* f_row, int n_fun0,
                                          double vol, int mbsize, void* vmat, void* workspace, size_t ws_bytes,
                                          void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (!gindex || !g_start || !glen0 || !pos || !sing || !f_sing || !f_atom || !f_row || !vmat || n_g <= 0 || n_sing <= 0 ||
      n_short <= 0 || ndf <= 0 || n_fun0 <= 0 || !(vol > 0.0) || mbsize <= 0 || (long)ndf * ndf * n_short > (1L << 31))
    return PYGW_E_ARG;
  size_t need = 0;
  pygw_zero_momentum_workspace(ndf, n_short, &need);
  void* ws = workspace;
  bool own = false;
  if (!ws) {
    HIP_TRY(hipMallocAsync(&ws, need, stream));
    own = true;
  } else if (ws_bytes < need) {
    return PYGW_E_ARG;
  }
  const long np = (long)ndf * ndf * n_short;
  hipLaunchKernelGGL(pygw_zm_phase, dim3(ceil_div(np, 256)), dim3(256), 0, stream, gindex, n_g, g_start, n_short, pos,
                     ndf, (double2*)ws);
  hipLaunchKernelGGL(pygw_zm_fill, dim3(ceil_div((long)n_fun0 * n_fun0, 4)), dim3(256), 0, stream, f_sing, f_atom,
                     f_row, n_fun0, sing, n_sing, glen0, n_short, (const double2*)ws, ndf, vol, mbsize, (double2*)vmat);
  int rc = hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
  if (own && hipFreeAsync(ws, stream) != hipSuccess && rc == PYGW_OK) rc = PYGW_E_HIP;
  return rc;
}
