// The LAPW augmentation coefficients and their band-basis products on gfx950.
//
//   gap2_set_lapwcoef (lapwcoeff.f90:3-222, called at pygw.py:2571,2595,2857):
//   for every plane wave or local-orbital row ig of a k-point and every atom
//   idf (type iat),
//     comc = i^l (4 pi R^2/sqrt(Vol)) e^{2 pi i (G + k).pos} conj(Y_lm(R (k+G)))
//     (conjugated when iop != 1),
//     alm = al_l(|k+G|) comc, blm = bl_l comc           (ig < nvk)
//     al = |k+G| j_l' pe - j_l dpe, bl = j_l dp - |k+G| j_l' p   (LAPW l)
//     al = j_l / p / R^2,          bl = 0                        (APW l)
//   at x = R |k+G|; a local-orbital row of (iat, l, lo) (lohns,
//   lapwcoeff.f90:421-449: (2l+1) mult rows per local orbital, atoms and l in
//   order) holds alo/blo/clo comc in the l block of every atom of type iat.
//   One thread per (row, atom, m >= 0): the Bessel functions of all l, then
//   Y_l,+-m for l >= m by the normalised associated-Legendre recurrence in l,
//   every alm/blm/clm entry written as it is produced; consecutive threads
//   hold consecutive rows, so each store of a wave is one contiguous run of
//   the Fortran-ordered arrays.
//
//   alfa = Aeig alm, beta = Aeig blm, gama = Aeig clm (pygw.py:2577-2583):
//   one complex GEMM launch on FP64 MFMA for the three (z = 0, 1, 2), the
//   coefficient arrays read as [ngk, columns] column-major.
#include "common.h"

typedef double d4 __attribute__((ext_vector_type(4)));

#include <algorithm>
#include <cmath>

namespace lw {
constexpr int kMaxNt = 16;  // lmax + 1
constexpr int kHdr = 16;    // plan header doubles

// plan (doubles): header [nat, nt, lomax, nloat, ndf, nLOmax, vol, off_df,
// atom block size, df block size, 0...], k2cartes at [kHdr, kHdr + 9)
// (Fortran order), then per atom type: rmt, mult, first idf, lapw[nt],
// nlo[lomax+1], nLO_at[lomax+1], umt[nt][4], abcelo[lomax+1][nloat][3],
// rotloc[9] (Fortran order); per atom: iat, pos[3], trotij[9].
enum { H_NAT, H_NT, H_LOMAX, H_NLOAT, H_NDF, H_NLOMAX, H_VOL, H_OFFDF, H_ATW, H_DFW };

__host__ __device__ inline int at_w(int nt, int lomax, int nloat) {
  return 3 + nt + 2 * (lomax + 1) + 4 * nt + 3 * (lomax + 1) * nloat + 9;
}
constexpr int kDfW = 13;

struct AtomView {
  const double* p;
  int nt, lomax, nloat;
  __host__ __device__ double rmt() const { return p[0]; }
  __host__ __device__ int mult() const { return (int)p[1]; }
  __host__ __device__ int idf0() const { return (int)p[2]; }
  __host__ __device__ int lapw(int l) const { return (int)p[3 + l]; }
  __host__ __device__ int nlo(int l) const { return (int)p[3 + nt + l]; }
  __host__ __device__ int nlo_at(int l) const { return (int)p[3 + nt + (lomax + 1) + l]; }
  __host__ __device__ double umt(int l, int j) const { return p[3 + nt + 2 * (lomax + 1) + 4 * l + j]; }
  __host__ __device__ double abcelo(int l, int jlo, int j) const {
    return p[3 + nt + 2 * (lomax + 1) + 4 * nt + 3 * (l * nloat + jlo) + j];
  }
  __host__ __device__ double rotloc(int i, int j) const {
    return p[3 + nt + 2 * (lomax + 1) + 4 * nt + 3 * (lomax + 1) * nloat + i + 3 * j];
  }
};

__host__ __device__ inline AtomView atom(const double* plan, int iat) {
  const int nt = (int)plan[H_NT], lomax = (int)plan[H_LOMAX], nloat = (int)plan[H_NLOAT];
  return AtomView{plan + kHdr + 9 + iat * at_w(nt, lomax, nloat), nt, lomax, nloat};
}

// j[l] = j_l(x), l = 0..L (L < kMaxNt + 1): the power series below x = 0.01;
// the upward recurrence from j_0 = sin x / x, j_1 = sin x / x^2 - cos x / x
// for x > L; otherwise Miller's downward recurrence from order L + 40 + x,
// normalised by sum_l (2l+1) j_l^2 = 1 with the sign of sin x / x or cos x / x
// (whichever is larger) against the recurrence's own j_0 or j_-1.
__device__ void bessel_all(int L, double x, double* j) {
  if (x < 0.01) {
    const double y = 0.5 * x * x;
    double xl_df = 1.0;  // x^l / (2l+1)!!
    for (int l = 0; l <= L; l++) {
      if (l > 0) xl_df *= x / (2.0 * l + 1.0);
      double term = 1.0, sum = 1.0;
      for (int k = 1; k <= 4; k++) {
        term *= -y / (k * (2.0 * l + 2.0 * k + 1.0));
        sum += term;
      }
      j[l] = xl_df * sum;
    }
    return;
  }
  double sx, cx;
  sincos(x, &sx, &cx);
  if (x > (double)L) {
    j[0] = sx / x;
    if (L >= 1) j[1] = (sx / x - cx) / x;
    for (int l = 1; l < L; l++) j[l + 1] = (2.0 * l + 1.0) / x * j[l] - j[l - 1];
    return;
  }
  const int N = L + 40 + (int)x;
  const double ix = 1.0 / x;
  double jp = 0.0, jc = 1e-100, norm = 0.0;
  for (int l = N; l >= 0; l--) {
    if (l <= L) j[l] = jc;
    norm += (2.0 * l + 1.0) * jc * jc;
    const double jm = (2.0 * l + 1.0) * ix * jc - jp;
    jp = jc;
    jc = jm;
    if (fabs(jc) > 1e150) {
      jc *= 1e-150;
      jp *= 1e-150;
      norm *= 1e-300;
      for (int i = l; i <= L; i++) j[i] *= 1e-150;
    }
  }
  // jp = j_0, jc = j_-1 (unnormalised)
  const double sgn = fabs(sx) >= fabs(cx) ? (jp * sx < 0.0 ? -1.0 : 1.0) : (jc * cx < 0.0 ? -1.0 : 1.0);
  const double s = sgn / sqrt(norm);
  for (int l = 0; l <= L; l++) j[l] *= s;
}

// the local orbital that row r (0-based, after the plane waves) belongs to:
// atom type, l and the 1-based orbital index, or iat = -1
__device__ inline void lo_row(const double* plan, int r, int& iat, int& l, int& jlo) {
  const int nat = (int)plan[H_NAT], lomax = (int)plan[H_LOMAX];
  int start = 0;
  for (int a = 0; a < nat; a++) {
    const AtomView A = atom(plan, a);
    for (int ll = 0; ll <= lomax; ll++) {
      const int blk = (2 * ll + 1) * A.mult(), n = A.nlo(ll);
      if (r < start + n * blk) {
        iat = a;
        l = ll;
        jlo = (r - start) / blk + 1;
        return;
      }
      start += n * blk;
    }
  }
  iat = -1;
}
}  // namespace lw

// alm/blm [ngk][nt^2][ndf], clm [ngk][nLOmax][nt^2][ndf] (Fortran order)
__global__ void __launch_bounds__(64) pygw_lapw_coef_k(const double* __restrict__ plan, double kl0, double kl1,
                                                       double kl2, int iop, int ifrot, int nvk,
                                                       const int* __restrict__ indgk, int ngk,
                                                       const int* __restrict__ gindex, int ng,
                                                       double2* __restrict__ alm, double2* __restrict__ blm,
                                                       double2* __restrict__ clm) {
  using namespace lw;
  const int ig = blockIdx.x * blockDim.x + threadIdx.x, idf = blockIdx.y, m = blockIdx.z;
  if (ig >= ngk) return;
  const int nt = (int)plan[H_NT], nlomax = (int)plan[H_NLOMAX];
  const int lmax = nt - 1, ntnt = nt * nt;
  const double* df = plan + (int)plan[H_OFFDF] + kDfW * idf;
  const int iat = (int)df[0];
  const AtomView A = atom(plan, iat);
  const double* k2c = plan + kHdr;  // k2cartes(i, j) at i + 3 j
  const double pi = 3.14159265358979;
  const int gi = indgk[ig];
  if (gi < 0 || gi >= ng) {  // not a row of gindex: the row reads as NaN
    const double nan = __builtin_nan("");
    for (int l = m; l <= lmax; l++)
      for (int sg = -1; sg <= 1; sg += 2) {
        if (m == 0 && sg < 0) continue;
        const int ilm = l * l + l + sg * m;
        const size_t e = (size_t)ig + (size_t)ngk * ((size_t)idf * ntnt + ilm);
        alm[e] = blm[e] = double2{nan, nan};
        for (int s = 0; s < nlomax; s++)
          clm[(size_t)ig + (size_t)ngk * (s + (size_t)nlomax * ((size_t)idf * ntnt + ilm))] = double2{nan, nan};
      }
    return;
  }
  const double G[3] = {(double)gindex[gi], (double)gindex[gi + ng], (double)gindex[gi + 2 * ng]};
  const double kl[3] = {kl0, kl1, kl2};
  double kpg[3];
  for (int i = 0; i < 3; i++) {
    const double kc = k2c[i] * kl[0] + k2c[i + 3] * kl[1] + k2c[i + 6] * kl[2];
    const double gc = k2c[i] * G[0] + k2c[i + 3] * G[1] + k2c[i + 6] * G[2];
    kpg[i] = kc + gc;
  }
  const double rk = sqrt(kpg[0] * kpg[0] + kpg[1] * kpg[1] + kpg[2] * kpg[2]);
  const double arg = 2.0 * pi * ((G[0] + kl[0]) * df[1] + (G[1] + kl[1]) * df[2] + (G[2] + kl[2]) * df[3]);
  double sa, ca;
  sincos(arg, &sa, &ca);
  double v[3] = {kpg[0], kpg[1], kpg[2]};
  if (ifrot) {
    double r1[3];
    for (int j = 0; j < 3; j++) r1[j] = v[0] * df[4 + 3 * j] + v[1] * df[5 + 3 * j] + v[2] * df[6 + 3 * j];
    for (int i = 0; i < 3; i++) v[i] = A.rotloc(i, 0) * r1[0] + A.rotloc(i, 1) * r1[1] + A.rotloc(i, 2) * r1[2];
  }
  // radial factors of this row per l: f1 multiplies into alm, f2 into blm,
  // f3 into clm at slot lo3 (or none)
  double f1[kMaxNt], f2[kMaxNt];
  int lrow = -1, lo3 = -1;
  double f3 = 0.0;
  if (ig < nvk) {
    const double R = A.rmt(), x = R * rk;
    double j[kMaxNt + 1];
    bessel_all(nt, x, j);
    for (int l = 0; l < nt; l++) {
      double dj;
      if (x > 0.0)
        dj = (l / x) * j[l] - j[l + 1];
      else
        dj = l == 1 ? 1.0 / 3.0 : 0.0;
      const double p = A.umt(l, 0), pe = A.umt(l, 1), dp = A.umt(l, 2), dpe = A.umt(l, 3);
      if (A.lapw(l) == 1) {
        f1[l] = rk * dj * pe - j[l] * dpe;
        f2[l] = j[l] * dp - rk * dj * p;
      } else {
        f1[l] = j[l] / p / (R * R);
        f2[l] = 0.0;
      }
    }
  } else {
    int ra, rl, rj;
    lo_row(plan, ig - nvk, ra, rl, rj);
    if (ra == iat && rl <= lmax) {
      const int ilo0 = A.lapw(rl), jlo = rj + ilo0 - 1;
      if (jlo <= A.nlo_at(rl)) {
        lrow = rl;
        f1[0] = A.abcelo(rl, jlo, 0);
        f2[0] = A.abcelo(rl, jlo, 1);
        f3 = A.abcelo(rl, jlo, 2);
        lo3 = jlo >= 1 ? jlo - 1 : -1;
      }
    }
  }
  const double rho = 4.0 * pi * sqrt(1.0 / plan[H_VOL]) * A.rmt() * A.rmt();
  // e^{i arg} rho: the l-independent part of comc
  const double cr = rho * ca, ci = rho * sa;
  const double rxy = sqrt(v[0] * v[0] + v[1] * v[1]), r = sqrt(rxy * rxy + v[2] * v[2]);
  const double ct = r > 0.0 ? v[2] / r : 1.0, st = r > 0.0 ? rxy / r : 0.0;
  const double cp = rxy > 0.0 ? v[0] / rxy : 1.0, sp = rxy > 0.0 ? v[1] / rxy : 0.0;
  const size_t col = (size_t)ngk, base = (size_t)idf * ntnt;
  auto put = [&](int l, int ilm, double yr, double yi) {
    // comc = i^l e^{i arg} rho conj(Y)
    double zr = cr * yr + ci * yi, zi = ci * yr - cr * yi;
    for (int t = 0; t < (l & 3); t++) {
      const double tr = -zi;
      zi = zr;
      zr = tr;
    }
    if (iop != 1) zi = -zi;
    double a1 = 0.0, a2 = 0.0, a3 = 0.0;
    if (ig < nvk) {
      a1 = f1[l];
      a2 = f2[l];
    } else if (l == lrow) {
      a1 = f1[0];
      a2 = f2[0];
      a3 = f3;
    }
    const size_t e = (size_t)ig + col * (base + ilm);
    alm[e] = double2{a1 * zr, a1 * zi};
    blm[e] = double2{a2 * zr, a2 * zi};
    for (int s = 0; s < nlomax; s++) {
      const double c = s == lo3 ? a3 : 0.0;
      clm[(size_t)ig + col * (s + (size_t)nlomax * (base + ilm))] = double2{c * zr, c * zi};
    }
  };
  // P_mm and e^{i m phi} by the same products as a loop over m would form
  double pmm = 1.0 / sqrt(4.0 * 3.14159265358979323846);
  double cm = 1.0, sm = 0.0;
  for (int k = 1; k <= m; k++) {
    pmm *= -sqrt((2.0 * k + 1.0) / (2.0 * k)) * st;
    const double c2 = cm * cp - sm * sp, s2 = sm * cp + cm * sp;
    cm = c2;
    sm = s2;
  }
  double p2 = 0.0, p1 = pmm;
  for (int l = m; l <= lmax; l++) {
    double p;
    if (l == m) {
      p = pmm;
    } else if (l == m + 1) {
      p = sqrt(2.0 * m + 3.0) * ct * pmm;
    } else {
      const double a = sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
      const double b = sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
      p = a * (ct * p1 - b * p2);
    }
    if (l > m) {
      p2 = p1;
      p1 = p;
    }
    const double yr = p * cm, yi = p * sm;
    put(l, l * l + l + m, yr, yi);
    if (m > 0) {
      const double sg = (m & 1) ? -1.0 : 1.0;  // Y_l,-m = (-1)^m conj(Y_lm)
      put(l, l * l + l - m, sg * yr, -sg * yi);
    }
  }
}

// out[z][n + nb c] = sum_g A(n, g) X_z[g + ngk c] (A(n, g) = Aeig[n lda + g],
// conjugated with conj_a), columns c < ncol_z.  Tiles of 32 bands x 32
// columns (the band window is ~140 wide: 64-row tiles would leave a third of
// the last one empty and too few blocks to fill 256 CUs), 4 waves of 16 x 16
// on v_mfma_f64_16x16x4f64.  Both operands are contiguous in g, so the LDS
// staging walks g fastest for both.
__global__ void __launch_bounds__(256) pygw_lapw_band_k(const double2* __restrict__ Aeig, long lda, int conj_a,
                                                       int nb, int ngk, const double2* __restrict__ X0,
                                                       const double2* __restrict__ X1, const double2* __restrict__ X2,
                                                       long ncol01, long ncol2, double2* __restrict__ O0,
                                                       double2* __restrict__ O1, double2* __restrict__ O2) {
  constexpr int T = 32, KC = 32;
  __shared__ double2 As[T][KC + 1];  // [band][g]
  __shared__ double2 Bs[T][KC + 1];  // [column][g]
  const int z = blockIdx.z;
  const long ncol = z < 2 ? ncol01 : ncol2;
  const long n0 = (long)blockIdx.x * T;
  const int m0 = blockIdx.y * T;
  if (n0 >= ncol) return;
  const double2* X = z == 0 ? X0 : (z == 1 ? X1 : X2);
  double2* O = z == 0 ? O0 : (z == 1 ? O1 : O2);
  const double cs = conj_a ? -1.0 : 1.0;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1, li = lane & 15, lk = lane >> 4;
  d4 aR = d4{0, 0, 0, 0}, aI = d4{0, 0, 0, 0};
  for (int k0 = 0; k0 < ngk; k0 += KC) {
    for (int e = tid; e < T * KC; e += 256) {
      const int c = e % KC, m = e / KC, g = k0 + c;
      double2 a{0.0, 0.0}, b{0.0, 0.0};
      if (g < ngk) {
        if (m0 + m < nb) a = Aeig[(size_t)(m0 + m) * lda + g];
        if (n0 + m < ncol) b = X[(size_t)g + (size_t)ngk * (n0 + m)];
      }
      As[m][c] = double2{a.x, cs * a.y};
      Bs[m][c] = b;
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
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int r = m0 + 16 * wr + 4 * v + lk;
    const long c = n0 + 16 * wc + li;
    if (r < nb && c < ncol) O[(size_t)r + (size_t)nb * c] = double2{aR[v], aI[v]};
  }
}

extern "C" int pygw_lapw_plan(int nat, const int* mult, int nt, int lomax, int nloat, const int* lapw,
                              const int* nlo, const int* nLO_at, const double* umt, const double* abcelo,
                              const double* rmt, const double* pos, const double* rotloc, const double* trotij,
                              double vol, const double* k2cartes, double* plan, int* plan_doubles) {
  using namespace lw;
  if (nat <= 0 || !mult || nt <= 0 || nt > kMaxNt || lomax < 0 || nloat <= 0 || !lapw || !nlo || !nLO_at || !umt ||
      !abcelo || !rmt || !pos || !rotloc || !trotij || !k2cartes || !(vol > 0.0) || !plan_doubles)
    return PYGW_E_ARG;
  int ndf = 0, nlomax = 0;
  for (int a = 0; a < nat; a++) {
    if (mult[a] <= 0 || !(rmt[a] > 0.0)) return PYGW_E_ARG;
    ndf += mult[a];
    for (int l = 0; l < nt; l++) {
      if (lapw[l + nt * a] != 0 && lapw[l + nt * a] != 1) return PYGW_E_ARG;
      // APW l: al = j / p / R^2 needs p != 0
      if (lapw[l + nt * a] == 0 && umt[a + nat * l] == 0.0) return PYGW_E_ARG;
    }
    for (int l = 0; l <= lomax; l++) {
      const int n = nlo[l + (lomax + 1) * a], nl = nLO_at[l + (lomax + 1) * a];
      if (n < 0 || nl < 0 || nl + 1 > nloat) return PYGW_E_ARG;
      // the local orbitals of (l, atom) stay inside their own lohns block
      if (l < nt && nl > n - 1 + lapw[l + nt * a]) return PYGW_E_ARG;
      nlomax = std::max(nlomax, nl);
    }
  }
  const int atw = at_w(nt, lomax, nloat);
  const int off_df = kHdr + 9 + nat * atw, total = off_df + kDfW * ndf;
  if (plan) {
    if (*plan_doubles < total) return PYGW_E_ARG;
    std::fill(plan, plan + total, 0.0);
    plan[H_NAT] = nat;
    plan[H_NT] = nt;
    plan[H_LOMAX] = lomax;
    plan[H_NLOAT] = nloat;
    plan[H_NDF] = ndf;
    plan[H_NLOMAX] = nlomax;
    plan[H_VOL] = vol;
    plan[H_OFFDF] = off_df;
    plan[H_ATW] = atw;
    plan[H_DFW] = kDfW;
    for (int i = 0; i < 9; i++) plan[kHdr + i] = k2cartes[i];
    int idf = 0;
    for (int a = 0; a < nat; a++) {
      double* p = plan + kHdr + 9 + a * atw;
      p[0] = rmt[a];
      p[1] = mult[a];
      p[2] = idf;
      int o = 3;
      for (int l = 0; l < nt; l++) p[o++] = lapw[l + nt * a];
      for (int l = 0; l <= lomax; l++) p[o++] = nlo[l + (lomax + 1) * a];
      for (int l = 0; l <= lomax; l++) p[o++] = nLO_at[l + (lomax + 1) * a];
      // umt(nat, nt, 5) Fortran order: (a, l, j) at a + nat (l + nt j)
      for (int l = 0; l < nt; l++)
        for (int j = 0; j < 4; j++) p[o++] = umt[a + (size_t)nat * (l + (size_t)nt * j)];
      // abcelo(nat, lomax+1, nloat, 3)
      for (int l = 0; l <= lomax; l++)
        for (int jl = 0; jl < nloat; jl++)
          for (int j = 0; j < 3; j++)
            p[o++] = abcelo[a + (size_t)nat * (l + (size_t)(lomax + 1) * (jl + (size_t)nloat * j))];
      // rotloc(nat, 3, 3): (a, i, j) at a + nat (i + 3 j)
      for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) p[o++] = rotloc[a + (size_t)nat * (i + 3 * j)];
      for (int q = 0; q < mult[a]; q++, idf++) {
        double* d = plan + off_df + kDfW * idf;
        d[0] = a;
        for (int i = 0; i < 3; i++) d[1 + i] = pos[i + 3 * idf];
        // trotij(ndf, 3, 3): (idf, i, j) at idf + ndf (i + 3 j)
        for (int j = 0; j < 3; j++)
          for (int i = 0; i < 3; i++) d[4 + i + 3 * j] = trotij[idf + (size_t)ndf * (i + 3 * j)];
      }
    }
  }
  *plan_doubles = total;
  return PYGW_OK;
}

static bool lapw_plan_ok(const double* p) {
  using namespace lw;
  return p && p[H_NAT] >= 1 && p[H_NT] >= 1 && p[H_NT] <= kMaxNt && p[H_NDF] >= 1 && p[H_NLOMAX] >= 0 &&
         p[H_VOL] > 0.0 && (int)p[H_ATW] == at_w((int)p[H_NT], (int)p[H_LOMAX], (int)p[H_NLOAT]);
}

extern "C" int pygw_lapw_coef(const double* plan_host, const double* plan_dev, const double* kl, int iop, int ifrot,
                              int nvk, const int* indgk, int ngk, const int* gindex, int ng, void* alm, void* blm,
                              void* clm, void* stream) {
  using namespace lw;
  if (!lapw_plan_ok(plan_host) || !plan_dev || !kl || !indgk || !gindex || !alm || !blm || ngk <= 0 || nvk < 0 ||
      nvk > ngk || ng <= 0 || (plan_host[H_NLOMAX] > 0 && !clm))
    return PYGW_E_ARG;
  const int ndf = (int)plan_host[H_NDF];
  hipLaunchKernelGGL(pygw_lapw_coef_k, dim3(ceil_div(ngk, 64), ndf, (int)plan_host[H_NT]), dim3(64), 0,
                     (hipStream_t)stream, plan_dev,
                     kl[0], kl[1], kl[2], iop, ifrot ? 1 : 0, nvk, indgk, ngk, gindex, ng, (double2*)alm,
                     (double2*)blm, (double2*)clm);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

extern "C" int pygw_lapw_band(const void* Aeig, long lda, int conj_a, int nb, int ngk, const void* alm,
                              const void* blm, const void* clm, long ncol, int nlomax, void* alfa, void* beta,
                              void* gama, void* stream) {
  if (!Aeig || lda < ngk || nb <= 0 || ngk <= 0 || !alm || !blm || !alfa || !beta || ncol <= 0 || nlomax < 0 ||
      (nlomax > 0 && (!clm || !gama)))
    return PYGW_E_ARG;
  const long ncol2 = ncol * nlomax;
  const int zs = nlomax > 0 ? 3 : 2;
  hipLaunchKernelGGL(pygw_lapw_band_k, dim3(ceil_div(std::max(ncol, ncol2), 32), ceil_div(nb, 32), zs), dim3(256), 0,
                     (hipStream_t)stream, (const double2*)Aeig, lda, conj_a, nb, ngk, (const double2*)alm,
                     (const double2*)blm, (const double2*)clm, ncol, ncol2, (double2*)alfa, (double2*)beta,
                     (double2*)gama);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}
