// The muffin-tin / plane-wave blocks of the Coulomb matrix on gfx950:
//
//   Mixed_Coulomb (for_Coulomb.f90:327-408, called at pygw.py:2396): for the
//   product functions (irm, L, M) of the equivalent atoms of one atom,
//     vtemp[im][ig] = Y_LM(R (q + G_ig)) e^{-2 pi i pos . G_ig} (4 pi/sqrt(Vol))
//                     (-i)^L jlam[irm][ipw(ig)] (4 pi / gqlen0[ipw]^2 for V)
//   (ipw the equal-|q+G| group of ig, the G = 0 column dropped at q = 0);
//   the blocks Vmat[loct + i][im] = sum_ig mpwipw[i][ig] Vmatit[im][ig] and
//   Vmat[im][loct + i] = its conjugate (pygw.py:2409-2416).
//
// Y_lm (Condon-Shortley phase, index l*l + l + m, as ylm in sphbes.f90) from
// the normalised associated Legendre recurrences
//   P_mm = -sqrt((2m+1)/(2m)) sin(t) P_(m-1)(m-1),  P_(m+1)m = sqrt(2m+3) cos(t) P_mm,
//   P_lm = a_lm (cos(t) P_(l-1)m - b_lm P_(l-2)m),
//   a_lm = sqrt((4l^2-1)/(l^2-m^2)), b_lm = sqrt(((l-1)^2-m^2)/(4(l-1)^2-1)),
// times e^{i m phi}; Y_l,-m = (-1)^m conj(Y_lm).
#include "gemm_tile.h"

namespace mc {
constexpr int kMaxL = 8;      // L of the product functions
constexpr int kMaxEq = 24;    // equivalent atoms per call
constexpr int kMaxMix = 64;   // product functions per atom
constexpr int NLM = (kMaxL + 1) * (kMaxL + 1);

struct Atoms {
  double rot[kMaxEq][9];  // rotloc[iat] trotij[idf]^T k2cartes
  double pos[kMaxEq][3];  // fractional position of each equivalent atom
  int big_l[kMaxMix];
  double vq[3];
  int neq, nmix, per, lmax;
};

// y[l*l + l + m] = Y_lm(v) for l <= lmax (v = 0: the z axis, phi = 0)
__device__ inline void ylm(const double v[3], int lmax, double2* y) {
  const double pi = 3.14159265358979323846;
  const double rxy = sqrt(v[0] * v[0] + v[1] * v[1]), r = sqrt(rxy * rxy + v[2] * v[2]);
  const double ct = r > 0.0 ? v[2] / r : 1.0, st = r > 0.0 ? rxy / r : 0.0;
  const double cp = rxy > 0.0 ? v[0] / rxy : 1.0, sp = rxy > 0.0 ? v[1] / rxy : 0.0;
  double pmm = 1.0 / sqrt(4.0 * pi);
  double cm = 1.0, sm = 0.0;  // cos(m phi), sin(m phi)
  for (int m = 0; m <= lmax; m++) {
    if (m > 0) {
      pmm *= -sqrt((2.0 * m + 1.0) / (2.0 * m)) * st;
      const double c2 = cm * cp - sm * sp, s2 = sm * cp + cm * sp;
      cm = c2;
      sm = s2;
    }
    double p2 = 0.0, p1 = pmm;
    for (int l = m; l <= lmax; l++) {
      double p;
      if (l == m)
        p = pmm;
      else if (l == m + 1)
        p = sqrt(2.0 * m + 3.0) * ct * pmm;
      else {
        const double a = sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
        const double b = sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
        p = a * (ct * p1 - b * p2);
      }
      if (l > m) {
        p2 = p1;
        p1 = p;
      }
      const double2 yp{p * cm, p * sm};
      y[l * l + l + m] = yp;
      if (m > 0) {
        const double sg = (m & 1) ? -1.0 : 1.0;
        y[l * l + l - m] = double2{sg * yp.x, -sg * yp.y};
      }
    }
  }
}
}  // namespace mc

// sph[(ieq ngq_barc + ig) NLM + lm] = Y_lm(R_ieq (q + G_ig))
__global__ void __launch_bounds__(256) pygw_mc_ylm(mc::Atoms at, const int* __restrict__ gindex,
                                                  const int* __restrict__ indgq, int ngq_barc,
                                                  double2* __restrict__ sph) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= at.neq * ngq_barc) return;
  const int ieq = e / ngq_barc, ig = e % ngq_barc;
  const int* g = gindex + 3 * (size_t)indgq[ig];
  const double k[3] = {at.vq[0] + g[0], at.vq[1] + g[1], at.vq[2] + g[2]};
  const double* R = at.rot[ieq];
  const double v[3] = {R[0] * k[0] + R[1] * k[1] + R[2] * k[2], R[3] * k[0] + R[4] * k[1] + R[5] * k[2],
                       R[6] * k[0] + R[7] * k[1] + R[8] * k[2]};
  mc::ylm(v, at.lmax, sph + (size_t)e * mc::NLM);
}

// out[(row0 + im) ld + ig] = vtemp[im][ig]
__global__ void __launch_bounds__(256) pygw_mc_fill(mc::Atoms at, const double2* __restrict__ sph,
                                                   const int* __restrict__ gindex, const int* __restrict__ indgq,
                                                   int ngq_barc, const double* __restrict__ jlam,
                                                   const double* __restrict__ gqlen0,
                                                   const int* __restrict__ G_unique, int ngs, int q_vq, double vol,
                                                   double2* __restrict__ out, long ld, int row0) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)at.neq * at.per * ngq_barc) return;
  const int ig = (int)(e % ngq_barc), im = (int)(e / ngq_barc);
  const int ieq = im / at.per;
  int r = im % at.per, irm = 0;
  while (r >= 2 * at.big_l[irm] + 1) {
    r -= 2 * at.big_l[irm] + 1;
    irm++;
  }
  const int L = at.big_l[irm], M = r - L;
  double2 v{0.0, 0.0};
  const bool q0 = fabs(at.vq[0]) + fabs(at.vq[1]) + fabs(at.vq[2]) < 1e-7;
  // the group ipw of ig: the last start G_unique[ipw] <= ig
  int lo = 0, hi = ngs - 1, ipw = -1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    if (G_unique[mid] <= ig) {
      ipw = mid;
      lo = mid + 1;
    } else
      hi = mid - 1;
  }
  if (ipw >= 0 && !(q0 && ig == 0)) {
    const double pi = 3.14159265358979323846;
    double c = 4.0 * pi / sqrt(vol) * jlam[(size_t)irm * ngs + ipw];
    if (q_vq) c *= 4.0 * pi / (gqlen0[ipw] * gqlen0[ipw]);
    // (-i)^L
    const double2 il = (L & 3) == 0 ? double2{1, 0} : (L & 3) == 1 ? double2{0, -1} : (L & 3) == 2 ? double2{-1, 0}
                                                                                                   : double2{0, 1};
    const int* g = gindex + 3 * (size_t)indgq[ig];
    const double ph = -2.0 * pi * (at.pos[ieq][0] * g[0] + at.pos[ieq][1] * g[1] + at.pos[ieq][2] * g[2]);
    double sn, cs;
    sincos(ph, &sn, &cs);
    const double2 y = sph[((size_t)ieq * ngq_barc + ig) * mc::NLM + L * L + L + M];
    const double2 ye{y.x * cs - y.y * sn, y.x * sn + y.y * cs};
    v = double2{c * (ye.x * il.x - ye.y * il.y), c * (ye.x * il.y + ye.y * il.x)};
  }
  out[(size_t)(row0 + im) * ld + ig] = v;
}

// mat2[i][im] = sum_ig mp[i][ig] W[im][ig]; V[(loct + i) ldv + im] = mat2,
// V[im ldv + loct + i] = conj(mat2) (W = Vmatit [loct][ngq_barc])
__global__ void __launch_bounds__(256) pygw_mc_assemble(const double2* __restrict__ mp, int ngq, int ngq_barc,
                                                       const double2* __restrict__ W, int loct,
                                                       double2* __restrict__ V, long ldv) {
  using namespace is;
  __shared__ double2 As[64][KC + 1];
  __shared__ double2 Bs[KC][64 + 1];
  const int ntn = (loct + 63) / 64, m0 = (blockIdx.x / ntn) * 64, n0 = (blockIdx.x % ntn) * 64;
  d4 aR[2][2], aI[2][2];
  zero(aR, aI);
  auto la = [&](int m, int k) -> double2 {
    const int r = m0 + m;
    return (r < ngq && k < ngq_barc) ? mp[(size_t)r * ngq_barc + k] : double2{0.0, 0.0};
  };
  auto lb = [&](int k, int c) -> double2 {
    const int col = n0 + c;
    return (k < ngq_barc && col < loct) ? W[(size_t)col * ngq_barc + k] : double2{0.0, 0.0};
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
        const int i = m0 + 32 * wr + 16 * a + 4 * v + lk, im = n0 + 32 * wc + 16 * b + li;
        if (i < ngq && im < loct) {
          const double2 x{aR[a][b][v], aI[a][b][v]};
          V[(size_t)(loct + i) * ldv + im] = x;
          V[(size_t)im * ldv + loct + i] = double2{x.x, -x.y};
        }
      }
}

int mixed_coulomb_launch(const double* vq, int iat, int q_vq, const double* jlam, int ngs, const int* gindex,
                         const int* indgq, const double* gqlen0, const int* G_unique, int ngq_barc, const int* big_l,
                         int nmix, const double* k2cartes, const double* rotloc, const double* trotij, const int* mult,
                         const double* pos, int nat, double vol, void* out, long ld, int row0, void* sph_ws,
                         hipStream_t stream) {
  using namespace mc;
  if (iat < 0 || iat >= nat || nmix <= 0 || nmix > kMaxMix || mult[iat] <= 0 || mult[iat] > kMaxEq) return PYGW_E_ARG;
  Atoms at{};
  at.neq = mult[iat];
  at.nmix = nmix;
  at.per = 0;
  at.lmax = 0;
  for (int i = 0; i < nmix; i++) {
    if (big_l[i] < 0 || big_l[i] > kMaxL) return PYGW_E_ARG;
    at.big_l[i] = big_l[i];
    at.per += 2 * big_l[i] + 1;
    at.lmax = big_l[i] > at.lmax ? big_l[i] : at.lmax;
  }
  for (int c = 0; c < 3; c++) at.vq[c] = vq[c];
  int idf = 0;
  for (int i = 0; i < iat; i++) idf += mult[i];
  for (int e = 0; e < at.neq; e++, idf++) {
    // R = rotloc[iat] (trotij[idf]^T k2cartes)
    double t1[9];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        double s = 0.0;
        for (int k = 0; k < 3; k++) s += trotij[idf * 9 + k * 3 + i] * k2cartes[k * 3 + j];
        t1[i * 3 + j] = s;
      }
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        double s = 0.0;
        for (int k = 0; k < 3; k++) s += rotloc[iat * 9 + i * 3 + k] * t1[k * 3 + j];
        at.rot[e][i * 3 + j] = s;
      }
    for (int c = 0; c < 3; c++) at.pos[e][c] = pos[idf * 3 + c];
  }
  double2* sph = (double2*)sph_ws;
  hipLaunchKernelGGL(pygw_mc_ylm, dim3(ceil_div((long)at.neq * ngq_barc, 256)), dim3(256), 0, stream, at, gindex,
                     indgq, ngq_barc, sph);
  hipLaunchKernelGGL(pygw_mc_fill, dim3(ceil_div((long)at.neq * at.per * ngq_barc, 256)), dim3(256), 0, stream, at,
                     sph, gindex, indgq, ngq_barc, jlam, gqlen0, G_unique, ngs, q_vq, vol, (double2*)out, ld, row0);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

size_t mixed_coulomb_ws_bytes(int max_mult, int ngq_barc) {
  return (size_t)max_mult * ngq_barc * mc::NLM * 16;
}

int coul_mixed_block_launch(const void* mpwipw, int ngq, int ngq_barc, const void* vmatit, int loct, void* vmat,
                            long ld, hipStream_t stream) {
  hipLaunchKernelGGL(pygw_mc_assemble, dim3(ceil_div(ngq, 64) * ceil_div(loct, 64)), dim3(256), 0, stream,
                     (const double2*)mpwipw, ngq, ngq_barc, (const double2*)vmatit, loct, (double2*)vmat, ld);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}
