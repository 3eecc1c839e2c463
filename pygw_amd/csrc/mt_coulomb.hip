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
#include <cmath>

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

// Y_lm(v) of one (l, m) (Condon-Shortley phase, Y_l,-m = (-1)^m conj(Y_lm)):
// the normalised associated Legendre recurrences
//   P_mm = -sqrt((2m+1)/(2m)) sin(t) P_(m-1)(m-1),  P_(m+1)m = sqrt(2m+3) cos(t) P_mm,
//   P_lm = a_lm (cos(t) P_(l-1)m - b_lm P_(l-2)m),
// times e^{i m phi}; v = 0 and v on the z axis take phi = 0.
__device__ double2 ylm_one(const double v[3], int l, int m) {
  const int am = abs(m);
  const double rxy = sqrt(v[0] * v[0] + v[1] * v[1]), r = sqrt(rxy * rxy + v[2] * v[2]);
  const double ct = r > 0.0 ? v[2] / r : 1.0, st = r > 0.0 ? rxy / r : 0.0;
  const double cp = rxy > 0.0 ? v[0] / rxy : 1.0, sp = rxy > 0.0 ? v[1] / rxy : 0.0;
  double pmm = 1.0 / sqrt(4.0 * mt::kPi), cm = 1.0, sm = 0.0;
  for (int k = 1; k <= am; k++) {
    pmm *= -sqrt((2.0 * k + 1.0) / (2.0 * k)) * st;
    const double c2 = cm * cp - sm * sp, s2 = sm * cp + cm * sp;
    cm = c2;
    sm = s2;
  }
  double p = pmm;
  if (l > am) {
    double p2 = pmm, p1 = sqrt(2.0 * am + 3.0) * ct * pmm;
    for (int ll = am + 2; ll <= l; ll++) {
      const double a = sqrt((4.0 * ll * ll - 1.0) / ((double)ll * ll - (double)am * am));
      const double bb = sqrt(((ll - 1.0) * (ll - 1.0) - (double)am * am) / (4.0 * (ll - 1.0) * (ll - 1.0) - 1.0));
      const double pn = a * (ct * p1 - bb * p2);
      p2 = p1;
      p1 = pn;
    }
    p = p1;
  }
  if (m >= 0) return double2{p * cm, p * sm};
  const double sg = (am & 1) ? -1.0 : 1.0;
  return double2{sg * p * cm, -sg * p * sm};
}

// sgm[lm + nlm*pair] for the atom pair `pair` (i <= j, pair = i + j(j+1)/2) and
// lm = l*l + l + m: one thread per entry sums the real-space terms, then the
// reciprocal-space terms, over the candidate lattice points (i1, i2, i3) in
// [-imax, imax]^3 (i1 slowest; imax = int(cutoff / shortest basis vector) + 1
// as in genr_str_size) in that fixed order, and divides by Gamma(l + 1/2).
// The threads of a wave share the pair and so the sequence of points.
__global__ void __launch_bounds__(64) pygw_ew_sgm(const double* __restrict__ pq, int lam, const double* __restrict__ pos,
                                                  const double* __restrict__ br2, const double* __restrict__ rbas,
                                                  const double* __restrict__ alat, double vol, int ortho, double r_cf,
                                                  double g_cf, double eta, double2* __restrict__ sgm) {
  using namespace mt;
  const int nlm = (lam + 1) * (lam + 1);
  const int pair = blockIdx.x, lm = blockIdx.y * 64 + threadIdx.x;
  if (lm >= nlm) return;
  int j = 0;
  while ((j + 1) * (j + 2) / 2 <= pair) j++;
  const int i = pair - j * (j + 1) / 2;
  int l = 0;
  while ((l + 1) * (l + 1) <= lm) l++;
  const int m = lm - l * l - l;
  // r_ij in cartesian coordinates (rbas rows are the lattice vectors; arrays Fortran)
  double d[3], raa[3];
  for (int x = 0; x < 3; x++) d[x] = pos[x + 3 * i] - pos[x + 3 * j];
  for (int x = 0; x < 3; x++)
    raa[x] = ortho ? d[x] * alat[x] : d[0] * rbas[0 + 3 * x] + d[1] * rbas[1 + 3 * x] + d[2] * rbas[2 + 3 * x];
  double lr = 1e300, lg = 1e300;
  for (int k = 0; k < 3; k++) {
    double vr = 0.0, vg = 0.0;
    for (int x = 0; x < 3; x++) {
      vr += rbas[k + 3 * x] * rbas[k + 3 * x];
      vg += br2[x + 3 * k] * br2[x + 3 * k];
    }
    lr = fmin(lr, sqrt(vr));
    lg = fmin(lg, sqrt(vg));
  }
  // at most 129^3 candidate points per sum: a degenerate lattice gives NaN
  if (!(lr > 0.0) || !(lg > 0.0) || !(r_cf / lr < kMaxCells) || !(g_cf / lg < kMaxCells)) {
    sgm[lm + (size_t)nlm * pair] = double2{__builtin_nan(""), __builtin_nan("")};
    return;
  }
  double2 s1{0.0, 0.0}, s2{0.0, 0.0};
  {
    const int imax = (int)(r_cf / lr) + 1;
    for (int i1 = -imax; i1 <= imax; i1++)
      for (int i2 = -imax; i2 <= imax; i2++)
        for (int i3 = -imax; i3 <= imax; i3++) {
          double rp[3];
          for (int x = 0; x < 3; x++) rp[x] = rbas[0 + 3 * x] * i1 + rbas[1 + 3 * x] * i2 + rbas[2 + 3 * x] * i3 + raa[x];
          const double rl = sqrt(rp[0] * rp[0] + rp[1] * rp[1] + rp[2] * rp[2]);
          if (!(rl <= r_cf && rl > 1e-6)) continue;
          // Gamma_l(r) by Gamma_l = Gamma_(l-1) (l - 1/2)/r + r^(l-2) e^{-r^2/eta^2} / eta^(2l-1)
          const double gausr = exp(-rl * rl / (eta * eta));
          double gam = sqrt(kPiF) * erfc(rl / eta) / rl, rpow = 1.0 / rl, epow = eta;
          for (int ll = 1; ll <= l; ll++) {
            gam = gam * (ll - 0.5) / rl + rpow * gausr / epow;
            rpow *= rl;
            epow *= eta * eta;
          }
          double sq, cq;
          sincos(-(rp[0] * pq[0] + rp[1] * pq[1] + rp[2] * pq[2]), &sq, &cq);
          const double2 y = ylm_one(rp, l, m);
          const double2 c{gam * cq, gam * sq};
          s1.x += c.x * y.x - c.y * y.y;
          s1.y += c.x * y.y + c.y * y.x;
        }
  }
  {
    const int imax = (int)(g_cf / lg) + 1;
    const double pref = 4.0 * kPiF * sqrt(kPiF) / vol;
    // (-i)^l
    const double2 il = (l & 3) == 0 ? double2{1.0, 0.0} : (l & 3) == 1 ? double2{0.0, -1.0}
                     : (l & 3) == 2 ? double2{-1.0, 0.0} : double2{0.0, 1.0};
    for (int i1 = -imax; i1 <= imax; i1++)
      for (int i2 = -imax; i2 <= imax; i2++)
        for (int i3 = -imax; i3 <= imax; i3++) {
          double g[3], gpq[3];
          for (int x = 0; x < 3; x++) {
            g[x] = br2[x + 3 * 0] * i1 + br2[x + 3 * 1] * i2 + br2[x + 3 * 2] * i3;  // br2 columns: reciprocal vectors
            gpq[x] = g[x] + pq[x];
          }
          const double gl = sqrt(gpq[0] * gpq[0] + gpq[1] * gpq[1] + gpq[2] * gpq[2]);
          if (!(gl <= g_cf && gl > 1e-6)) continue;
          double sg, cg;
          sincos(g[0] * raa[0] + g[1] * raa[1] + g[2] * raa[2], &sg, &cg);
          // pref e^{-eta^2 gl^2/4} gl^-2 (gl/2)^l (-i)^l e^{i G.r_ij}
          double a = pref * exp(-(eta * gl) * (eta * gl) / 4.0) / (gl * gl);
          for (int ll = 1; ll <= l; ll++) a *= 0.5 * gl;
          const double2 ph{a * (il.x * cg - il.y * sg), a * (il.x * sg + il.y * cg)};
          const double2 y = ylm_one(gpq, l, m);
          s2.x += ph.x * y.x - ph.y * y.y;
          s2.y += ph.x * y.y + ph.y * y.x;
        }
  }
  double gam = sqrt(kPiF);
  for (int k = 1; k <= l; k++) gam *= (k - 0.5);
  double2 v{(s1.x + s2.x) / gam, (s1.y + s2.y) / gam};
  if (lm == 0 && i == j) v.x -= 1.0 / (eta * kPiF);
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
  double jp = 0.0, jc = 1e-100, norm = 0.0, want = 0.0;
  for (int l = N; l >= 0; l--) {
    if (l == L) want = jc;
    norm += (2.0 * l + 1.0) * jc * jc;
    const double jm = (2.0 * l + 1.0) / x * jc - jp;
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


// tab[(L*npt + j)*nk + k] = j_L(k r_j) r_j (mode 0, L <= lmax) or sin(k r_j)
// (mode 1, L = 0): one thread per entry, consecutive threads along k.
__global__ void __launch_bounds__(256) pygw_rad_table(int mode, int nl, int npt, const double* __restrict__ rx,
                                                      const double* __restrict__ k, int nk, double* __restrict__ tab) {
  const long e = blockIdx.x * 256L + threadIdx.x;
  if (e >= (long)nl * npt * nk) return;
  const int ik = (int)(e % nk);
  const long lj = e / nk;
  const int j = (int)(lj % npt), L = (int)(lj / npt);
  const double r = rx[j], x = k[ik] * r;
  tab[e] = mode == 0 ? sph_bessel(L, x) * r : sin(x);
}

// out[f][k] = sum_j w_j (A_f(j) + c2 B_f(j)) tab[L_f][j][k]: one thread per
// (f, k), the threads of a block share f (blockIdx.y).  L_f outside [0, nl)
// gives NaN.
//
// w_j is RINT13g's rule (radials.f90:140-187) for npt >= 4 points: Simpson's
// weights dh r_j (1, 4, 2, ..., 4, 1)/3 from the second point (1-based, odd
// npt) or the third (even npt, with a trapezoid dh (f_1 + f_2)/2 over the
// first interval), and r0 f(r0)/3 for [0, r0].  The r_j of the weights are
// the ones RINT13g forms: r0 D^(J0-1) at the first Simpson point J0 and a
// product by D per point after it (which drifts from r0 D^j by up to ~1e-13
// at the end of an 800-point mesh), r0 D^(J0-1) / D at point J0 - 1.
__global__ void __launch_bounds__(256) pygw_rad_olap(const double* __restrict__ ul, const double* __restrict__ us,
                                                     const int* __restrict__ lf, int nl, int npt,
                                                     const double* __restrict__ rx, double dh, double D, double c2,
                                                     const double* __restrict__ tab, int nk,
                                                     double* __restrict__ out) {
  const int f = blockIdx.y, ik = blockIdx.x * 256 + threadIdx.x;
  if (ik >= nk) return;
  const int L = lf[f];
  if (L < 0 || L >= nl) {
    out[(size_t)f * nk + ik] = __builtin_nan("");
    return;
  }
  const double* t = tab + (size_t)L * npt * nk + ik;
  const double* a = ul + (size_t)f * npt;
  const double* b = us + (size_t)f * npt;
  const double r0 = rx[0];
  const bool odd = (npt & 1) != 0;
  const int j0 = odd ? 1 : 2;                // 0-based first Simpson point
  double R = odd ? r0 * D : r0 * (D * D);     // its r
  const double R1 = R / D;                    // r of point j0 - 1
  double s = 0.0;
  for (int j = 0; j < npt; j++) {
    double w;
    if (j < j0) {
      if (j == 0)
        w = r0 / 3.0 + (odd ? dh * R1 / 3.0 : 0.5 * dh * r0);
      else  // even npt, second point
        w = dh * R1 / 3.0 + 0.5 * dh * R1;
    } else {
      w = (j == npt - 1 ? 1.0 : (((j - j0) & 1) == 0 ? 4.0 : 2.0)) * dh * R / 3.0;
      R *= D;
    }
    s += w * (a[j] + c2 * b[j]) * t[(size_t)j * nk];
  }
  out[(size_t)f * nk + ik] = s;
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
__global__ void __launch_bounds__(256) pygw_zm_fill(const int* __restrict__ fidx, const int* __restrict__ fdf,
                                                    const int* __restrict__ frow, int nf0, const double* __restrict__ sing,
                                                    int nsing,
                                                    const double* __restrict__ glen0, int ngs,
                                                    const double2* __restrict__ phase, int ndf, double vol, int mbs,
                                                    double2* __restrict__ V) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= nf0 * nf0) return;
  const int a = e % nf0, b = e / nf0;
  if (frow[a] < 0 || frow[a] >= mbs || frow[b] < 0 || frow[b] >= mbs) return;
  if (fidx[a] < 0 || fidx[a] >= nsing || fidx[b] < 0 || fidx[b] >= nsing || fdf[a] < 0 || fdf[a] >= ndf ||
      fdf[b] < 0 || fdf[b] >= ndf) {
    V[frow[a] + (size_t)mbs * frow[b]] = double2{__builtin_nan(""), __builtin_nan("")};
    return;
  }
  const double c = 16.0 * mt::kPi * mt::kPi / vol;
  const double* sa = sing + (size_t)fidx[a] * ngs;
  const double* sb = sing + (size_t)fidx[b] * ngs;
  const double2* ph = phase + ((size_t)fdf[a] * ndf + fdf[b]) * ngs;
  double2 s{0.0, 0.0};
  for (int g = 0; g < ngs; g++) {
    const double gl2 = glen0[g] * glen0[g];
    const double w = c / (gl2 * gl2) * sa[g] * sb[g];
    s.x += ph[g].x * w;
    s.y += ph[g].y * w;
  }
  V[frow[a] + (size_t)mbs * frow[b]] = s;
}

__global__ void pygw_zero_v2(double2* __restrict__ x, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    x[e] = double2{0.0, 0.0};
}

// ---------------------------------------------------------------------------
extern "C" int pygw_ewald_summation(const double* pq, int lam_max, const double* pos, const double* br2,
                                    const double* rbas, const double* alat, double vol, int ortho, double r_cf,
                                    double g_cf, double eta, int ndf, void* sgm, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (!pq || !pos || !br2 || !rbas || !alat || !sgm || lam_max < 0 || lam_max > mt::kMaxLam || ndf <= 0 ||
      ndf > 4096 || !(vol > 0.0) || !(r_cf > 0.0) || !(g_cf > 0.0) || !(eta > 0.0) || (ortho != 0 && ortho != 1))
    return PYGW_E_ARG;
  const int npair = ndf * (ndf + 1) / 2, nlm = (lam_max + 1) * (lam_max + 1);
  hipLaunchKernelGGL(pygw_ew_sgm, dim3(npair, ceil_div(nlm, 64)), dim3(64), 0, stream, pq, lam_max, pos, br2, rbas, alat,
                     vol, ortho, r_cf, g_cf, eta, (double2*)sgm);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
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
  *bytes = (size_t)(lmax + 1) * npt * n_k * sizeof(double);
  return PYGW_OK;
}

extern "C" int pygw_radial_olap(int mode, int rel, const double* ul, const double* us, const int* lf, int n_fun,
                                int lmax, int npt, const double* rx, double dh, const double* k, int n_k, double* out,
                                void* workspace, size_t ws_bytes, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  size_t need = 0;
  if (pygw_radial_olap_workspace(mode, lmax, npt, n_k, &need) != PYGW_OK || !ul || !us || !lf || !rx || !k || !out ||
      n_fun <= 0 || n_fun > 65535 || !(dh > 0.0) || (rel != 0 && rel != 1))
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
  const long nt = (long)(lmax + 1) * npt * n_k;
  // RINT13g's ratio e^dh from the host's libm, as the compiled reference
  // forms it (a last-bit difference in it grows j-fold along the mesh)
  const double D = std::exp(dh);
  hipLaunchKernelGGL(pygw_rad_table, dim3(ceil_div(nt, 256)), dim3(256), 0, stream, mode, lmax + 1, npt, rx, k, n_k,
                     (double*)ws);
  hipLaunchKernelGGL(pygw_rad_olap, dim3(ceil_div(n_k, 256), n_fun), dim3(256), 0, stream, ul, us, lf, lmax + 1, npt,
                     rx, dh, D, c2, (const double*)ws, n_k, out);
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
  hipLaunchKernelGGL(pygw_zm_fill, dim3(ceil_div((long)n_fun0 * n_fun0, 256)), dim3(256), 0, stream, f_sing, f_atom,
                     f_row, n_fun0, sing, n_sing, glen0, n_short, (const double2*)ws, ndf, vol, mbsize, (double2*)vmat);
  int rc = hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
  if (own && hipFreeAsync(ws, stream) != hipSuccess && rc == PYGW_OK) rc = PYGW_E_HIP;
  return rc;
}
