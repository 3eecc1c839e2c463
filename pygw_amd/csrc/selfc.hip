// Sigma_c tail on gfx950, FP64: the G0 (x) W frequency convolution
// (Compute_selfc_inside, mcommon.py:80-165, with fnc.f90:86-176), its
// cubic-spline refinement onto the precise mesh (mcommon.py:126-140), and
// the Pade continuation plus the quasiparticle equation
// (AnalyticContinuation iop_ac=0, mcommon.py:315-343; Compute_quasiparticles,
// mcommon.py:167-210; Padecof/PadeG, for_pade.f90:3-51).
#include <algorithm>

#include "common.h"

namespace sc {
constexpr double PI_F = 3.14159265358979;  // pi as written in fnc.f90

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return double2{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return double2{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
  if (fabs(b.x) >= fabs(b.y)) {
    double r = b.y / b.x, d = b.x + b.y * r;
    return double2{(a.x + a.y * r) / d, (a.y - a.x * r) / d};
  }
  double r = b.x / b.y, d = b.x * r + b.y;
  return double2{(a.x * r + a.y) / d, (a.y * r - a.x) / d};
}
__device__ __forceinline__ double2 clog_(double2 z) { return double2{log(hypot(z.x, z.y)), atan2(z.y, z.x)}; }
// principal atan(z) = (i/2) [log(1 - i z) - log(1 + i z)]
__device__ __forceinline__ double2 catan_(double2 z) {
  const double2 iz = double2{-z.y, z.x};
  const double2 l1 = clog_(double2{1.0 - iz.x, -iz.y});
  const double2 l2 = clog_(double2{1.0 + iz.x, iz.y});
  const double2 d = csub(l1, l2);
  return double2{-0.5 * d.y, 0.5 * d.x};
}
}  // namespace sc

// Cubic spline of y(x) = (x^2+1) * mwm(x) on the coarse mesh, bc y''(x_0)=0,
// y'(x_{n-1})=0 (scipy CubicSpline bc_type=((2,0),(1,0)), extrapolating with
// the end polynomials), evaluated on the precise mesh and divided by
// (xp^2+1).  The imaginary part is splined only where sum|Im|/n > 1e-10, as
// at mcommon.py:134-138.  Row = (k, ie1): mwm is PyGW's per-k mwm[iom, ie1,
// ie2] stacked over k, [N_k][nom][nb1][nb2]; out is [nrows][nb2][nomp].
// One block per (row, ie2) pair: the pair's coarse values into LDS, the
// two Thomas sweeps for the slopes (real, imaginary part) by two threads,
// then every thread evaluates precise points j = t, t + 128, ... so that the
// block writes its nomp values contiguously (one thread per pair wrote 64
// rows nomp*16 bytes apart at each store: 1.66 ms against 0.4 ms at C5).
__global__ void __launch_bounds__(128) pygw_selfc_spline_b(const double2* __restrict__ mwm, int nrows, int nb1,
                                                          int nom, int nb2, const double* __restrict__ x,
                                                          const double* __restrict__ xp, int nomp,
                                                          double2* __restrict__ out) {
  constexpr int NMAX = 128;
  __shared__ double yr[NMAX], yi[NMAX], sr[NMAX], si[NMAX], xs[NMAX], cpr[NMAX], cpi[NMAX];
  __shared__ double c2r[NMAX], c3r[NMAX], c2i[NMAX], c3i[NMAX];
  __shared__ double red[2];
  const long t0 = blockIdx.x;   // (row, ie2)
  const int row = (int)(t0 / nb2), ie2 = (int)(t0 % nb2);
  const int kq = row / nb1, ie1 = row % nb1;
  const int t = threadIdx.x, n = nom;
  double im = 0.0;
  for (int i = t; i < n; i += blockDim.x) {
    const double2 v = mwm[(((size_t)kq * nom + i) * nb1 + ie1) * nb2 + ie2];
    const double f = x[i] * x[i] + 1.0;
    yr[i] = f * v.x;
    yi[i] = f * v.y;
    xs[i] = x[i];
    im += fabs(v.y);
  }
  // sum |Im| over the coarse points (the order of the one-thread form's sum
  // does not matter for the > 1e-10 test up to rounding)
  for (int off = 32; off > 0; off >>= 1) im += __shfl_xor(im, off, 64);
  if ((t & 63) == 0) red[t >> 6] = im;
  __syncthreads();
  const bool doim = (red[0] + red[1]) / n > 1e-10;
  if (t < 2 && (t == 0 || doim)) {
    const double* y = t ? yi : yr;
    double* sl = t ? si : sr;
    double* cp = t ? cpi : cpr;     // in LDS: a private array indexed at run time would live in scratch
    const double b0 = 2.0, c0 = 1.0;
    const double d0 = 3.0 * (y[1] - y[0]) / (xs[1] - xs[0]);
    cp[0] = c0 / b0;
    sl[0] = d0 / b0;
    for (int i = 1; i < n - 1; i++) {
      const double h0 = xs[i] - xs[i - 1], h1 = xs[i + 1] - xs[i];
      const double m0 = (y[i] - y[i - 1]) / h0, m1 = (y[i + 1] - y[i]) / h1;
      const double a = h1, b = 2.0 * (h0 + h1), c = h0, d = 3.0 * (h1 * m0 + h0 * m1);
      const double den = b - a * cp[i - 1];
      cp[i] = c / den;
      sl[i] = (d - a * sl[i - 1]) / den;
    }
    sl[n - 1] = 0.0;
    for (int i = n - 2; i >= 0; i--) sl[i] = sl[i] - cp[i] * sl[i + 1];
  }
  __syncthreads();
  // the cubic of every coarse interval, once per pair
  for (int x2 = t; x2 < 2 * (n - 1); x2 += blockDim.x) {
    const int part = x2 / (n - 1), seg = x2 % (n - 1);
    if (part == 1 && !doim) continue;
    const double* y = part ? yi : yr;
    const double* sl = part ? si : sr;
    const double h = xs[seg + 1] - xs[seg];
    const double m = (y[seg + 1] - y[seg]) / h;
    (part ? c2i : c2r)[seg] = (3.0 * m - 2.0 * sl[seg] - sl[seg + 1]) / h;
    (part ? c3i : c3r)[seg] = (sl[seg] + sl[seg + 1] - 2.0 * m) / (h * h);
  }
  __syncthreads();
  double2* o = out + (size_t)t0 * nomp;
  for (int j = t; j < nomp; j += blockDim.x) {
    const double xv = xp[j];
    int lo = 0, hi = n - 2;            // the interval the one-thread form's scan stops at
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (xv > xs[mid + 1]) lo = mid + 1; else hi = mid;
    }
    const int seg = lo;
    const double tt = xv - xs[seg];
    double res[2] = {0.0, 0.0};
    res[0] = yr[seg] + tt * (sr[seg] + tt * (c2r[seg] + tt * c3r[seg]));
    if (doim) res[1] = yi[seg] + tt * (si[seg] + tt * (c2i[seg] + tt * c3i[seg]));
    const double f = xv * xv + 1.0;
    o[j] = double2{res[0] / f, res[1] / f};
  }
}

// sc[row][iom] = sum_ie2 fr_convolution2(enk[k][ie2], mwmp[row][ie2][:], mwm[row][iom][ie2],
//                                        omega_p, omega[iom], womeg_p)    (fnc.f90:86-119)
// row = (k, ie1); mwm = [N_k][nom][nb1][nb2]; mwmp = [nrows][nb2][nomp]; for
// the coarse form (all_fr_convolutions) pass mwmp = NULL and the coarse mesh.
// One block per (row, group of CG frequencies): the threads split the
// precise-mesh points l (coalesced reads of mwmp, each read once per group),
// sum over l per (ie2, iom) in registers, scale by eps_om/pi per ie2 as the
// reference does, then reduce over the threads.
namespace sc {
constexpr int CG = 8;   // Matsubara points per block
constexpr int CT = 256;  // threads per block
}  // namespace sc

__global__ void __launch_bounds__(256) pygw_selfc_conv(const double* __restrict__ enk, int nb1, int nb2,
                                                       const double2* __restrict__ mwm, int nom,
                                                       const double* __restrict__ omega,
                                                       const double2* __restrict__ mwmp, int nomp,
                                                       const double* __restrict__ omp, const double* __restrict__ womp,
                                                       double2* __restrict__ out) {
  using namespace sc;
  __shared__ double2 red[CT / 64][CG];
  const int row = blockIdx.x, i0 = blockIdx.y * CG;
  const int k = row / nb1, ie1 = row % nb1;
  const int t = threadIdx.x;
  const int ng = nom - i0 < CG ? nom - i0 : CG;
  const double wmax = omp[nomp - 1];
  double2 acc[CG];
#pragma unroll
  for (int g = 0; g < CG; g++) acc[g] = double2{0.0, 0.0};
  for (int ie2 = 0; ie2 < nb2; ie2++) {
    const double e = enk[(size_t)k * nb2 + ie2];
    double2 s0[CG];
#pragma unroll
    for (int g = 0; g < CG; g++) s0[g] = double2{0.0, 0.0};
    for (int l = t; l < nomp; l += CT) {
      const double2 ml = mwmp ? mwmp[((size_t)row * nb2 + ie2) * nomp + l]
                              : mwm[(((size_t)k * nom + l) * nb1 + ie1) * nb2 + ie2];
      const double wl = womp[l], o2 = omp[l] * omp[l];
#pragma unroll
      for (int g = 0; g < CG; g++) {
        if (g < ng) {
          const double wn = omega[i0 + g];
          const double2 mn = mwm[(((size_t)k * nom + i0 + g) * nb1 + ie1) * nb2 + ie2];
          const double2 num = double2{(ml.x - mn.x) * wl, (ml.y - mn.y) * wl};
          const double2 den = double2{o2 + (e * e - wn * wn), -2.0 * e * wn};  // omega_l^2 + (e - i w_n)^2
          // num / den as num * conj(den) / |den|^2 with 1/|den|^2 from the
          // hardware reciprocal and two Newton-Raphson steps (relative error
          // a few ulp; |den|^2 vanishes only where den does)
          const double d2 = den.x * den.x + den.y * den.y;
          double r = __builtin_amdgcn_rcp(d2);
          r = __builtin_fma(r, __builtin_fma(-d2, r, 1.0), r);
          r = __builtin_fma(r, __builtin_fma(-d2, r, 1.0), r);
          s0[g] = cadd(s0[g], double2{(num.x * den.x + num.y * den.y) * r, (num.y * den.x - num.x * den.y) * r});
        }
      }
    }
#pragma unroll
    for (int g = 0; g < CG; g++) {
      if (g < ng) {
        const double2 eo = double2{e, -omega[i0 + g]};
        const double2 r = cmul(eo, s0[g]);
        acc[g] = double2{acc[g].x + r.x / PI_F, acc[g].y + r.y / PI_F};
      }
    }
  }
  // the atan tail once per (ie2, iom), the ie2 dealt over the threads
  for (int ie2 = t; ie2 < nb2; ie2 += CT) {
    const double e = enk[(size_t)k * nb2 + ie2];
#pragma unroll
    for (int g = 0; g < CG; g++) {
      if (g < ng) {
        const double2 mn = mwm[(((size_t)k * nom + i0 + g) * nb1 + ie1) * nb2 + ie2];
        if (hypot(mn.x, mn.y) > 1e-16) {
          const double2 eo = double2{e, -omega[i0 + g]};
          const double2 at = catan_(cdiv(double2{wmax, 0.0}, eo));
          acc[g] = cadd(acc[g], cmul(double2{mn.x / PI_F, mn.y / PI_F}, at));
        }
      }
    }
  }
  const int lane = t & 63, wv = t >> 6;
#pragma unroll
  for (int g = 0; g < CG; g++) {
    double x = acc[g].x, y = acc[g].y;
    for (int off = 32; off > 0; off >>= 1) {
      x += __shfl_xor(x, off, 64);
      y += __shfl_xor(y, off, 64);
    }
    if (lane == 0) red[wv][g] = double2{x, y};
  }
  __syncthreads();
  if (t < ng) {
    double2 r = red[0][t];
    for (int w = 1; w < CT / 64; w++) r = cadd(r, red[w][t]);
    out[(size_t)row * nom + i0 + t] = r;
  }
}

// Pade (Thiele) through the first npts points of sigma(i w_n), evaluated at
// e and e +- 1e-3 (mcommon.py:330-338), then the quasiparticle update of
// Compute_quasiparticles (mcommon.py:182-203):
//   Z = 1/(1 - Re dS), set to 1 (and dS to 0) when Z > 1 or Z < 0;
//   s = Re S(e) + sigx;
//   iop_es = 0 (G0W0):        E_qp = e + Z (s - vxc)
//   iop_es = -1 (GW0), e0 the Kohn-Sham energy:
//     iop_gw0 = 1: E_qp = e + s - vxc + e0 - e
//     iop_gw0 = 2: E_qp = e + Z (s - vxc) + e0 - e
//     otherwise:   E_qp = e + Z (s - vxc + e0 - e)
// out per state: [S(e) re, im, dS re, im, Z, E_qp]
__global__ void pygw_pade_qp(const double2* __restrict__ sig, int nstate, int nom, const double* __restrict__ omega,
                             int npts, const double* __restrict__ ene, const double* __restrict__ ene0,
                             const double* __restrict__ sigx, const double* __restrict__ vxc, int iop_es, int iop_gw0,
                             double* __restrict__ out) {
  using namespace sc;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nstate) return;
  constexpr int NP = 64;
  double2 g[NP];
  const int n = npts;
  for (int j = 0; j < n; j++) g[j] = sig[(size_t)s * nom + j];
  // p(i,j) = (p(i-1,i-1) - p(i-1,j)) / (z_j - z_{i-1}) / p(i-1,j), row i in place
  for (int i = 1; i < n; i++) {
    const double2 pd = g[i - 1];
    for (int j = i; j < n; j++) {
      const double2 dz = double2{0.0, omega[j] - omega[i - 1]};
      g[j] = cdiv(cdiv(csub(pd, g[j]), dz), g[j]);
    }
  }
  const double e = ene[s];
  double2 val[3];
  const double xs[3] = {e, e + 1e-3, e - 1e-3};
  for (int q = 0; q < 3; q++) {
    const double2 xv = double2{xs[q], 0.0};
    double2 a0 = {0.0, 0.0}, a1 = g[0], b0 = {1.0, 0.0}, b1 = {1.0, 0.0};
    for (int i = 1; i < n; i++) {
      const double2 f = cmul(csub(xv, double2{0.0, omega[i - 1]}), g[i]);
      const double2 a2 = cadd(a1, cmul(f, a0));
      const double2 b2 = cadd(b1, cmul(f, b0));
      a0 = a1; a1 = a2; b0 = b1; b1 = b2;
    }
    val[q] = cdiv(a1, b1);
  }
  const double2 ds = double2{(val[1].x - val[2].x) / 2e-3, (val[1].y - val[2].y) / 2e-3};
  double z = 1.0 / (1.0 - ds.x);
  if (z > 1.0 || z < 0.0) z = 1.0;
  const double snk = val[0].x + (sigx ? sigx[s] : 0.0);
  const double vx = vxc ? vxc[s] : 0.0;
  double delta;
  if (iop_es == 0) {
    delta = z * (snk - vx);
  } else {
    const double e0 = ene0[s];
    if (iop_gw0 == 1)
      delta = snk - vx + e0 - e;
    else if (iop_gw0 == 2)
      delta = z * (snk - vx) + e0 - e;
    else
      delta = z * (snk - vx + e0 - e);
  }
  double* o = out + (size_t)s * 6;
  o[0] = val[0].x; o[1] = val[0].y; o[2] = ds.x; o[3] = ds.y; o[4] = z;
  o[5] = e + delta;
}

// Thiele reciprocal differences (Padecof, for_pade.f90:3-25) of n_state
// functions on common complex nodes zn[n]: coef[s][j] = p(j, j) with
//   p(1, j) = g(j),  p(i, j) = (p(i-1, i-1) - p(i-1, j)) / (zn(j) - zn(i-1)) / p(i-1, j),
// row i computed in place.  One thread per function; n <= 64.
__global__ void pygw_padecof(const double2* __restrict__ zn, const double2* __restrict__ gn, int n, int nstate,
                             double2* __restrict__ coef) {
  using namespace sc;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nstate) return;
  double2 g[64];
  for (int j = 0; j < n; j++) g[j] = gn[(size_t)s * n + j];
  for (int i = 1; i < n; i++) {
    const double2 pd = g[i - 1];
    for (int j = i; j < n; j++) g[j] = cdiv(cdiv(csub(pd, g[j]), csub(zn[j], zn[i - 1])), g[j]);
  }
  for (int j = 0; j < n; j++) coef[(size_t)s * n + j] = g[j];
}

// The continued fraction at x[m] (PadeG, for_pade.f90:30-51):
//   a(0) = 0, a(1) = P(1), b(0) = b(1) = 1,
//   a(i+1) = a(i) + (x - zn(i)) P(i+1) a(i-1)  (b likewise),  G(x) = a(n)/b(n).
// One thread per (function, point).
__global__ void pygw_padeg(const double2* __restrict__ x, int m, const double2* __restrict__ zn,
                           const double2* __restrict__ coef, int n, int nstate, double2* __restrict__ gx) {
  using namespace sc;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)nstate * m) return;
  const int s = (int)(t / m), j = (int)(t % m);
  const double2* P = coef + (size_t)s * n;
  const double2 xv = x[j];
  double2 a0 = {0.0, 0.0}, a1 = P[0], b0 = {1.0, 0.0}, b1 = {1.0, 0.0};
  for (int i = 1; i < n; i++) {
    const double2 f = cmul(csub(xv, zn[i - 1]), P[i]);
    const double2 a2 = cadd(a1, cmul(f, a0));
    const double2 b2 = cadd(b1, cmul(f, b0));
    a0 = a1; a1 = a2; b0 = b1; b1 = b2;
  }
  gx[t] = cdiv(a1, b1);
}

// out = beta * out + alpha * sum_q parts[q] (q in order): the sum over q of
// the per-q Sigma_c contributions of Compute_selfc (postprocess.py:176-186)
// and the self-energy mixing of the GW0 loop (postprocess.py:297).
__global__ void pygw_selfc_qsum(const double2* __restrict__ parts, int nq, long n, double alpha, double beta,
                                double2* __restrict__ out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double2 acc = parts[i];
    for (int q = 1; q < nq; q++) {
      const double2 v = parts[(size_t)q * n + i];
      acc.x += v.x;
      acc.y += v.y;
    }
    double2 o = double2{0.0, 0.0};
    if (beta != 0.0) o = out[i];
    out[i] = double2{beta * o.x + alpha * acc.x, beta * o.y + alpha * acc.y};
  }
}

// enk of each (q, irk) item of Compute_selfc_inside (mcommon.py:94-101): the
// core energies, then bands[jrk[item], 0 .. nb2 - ncore - 1]; bands is
// [n_irk][n_band] (row-major), enk [n_items][nb2].
__global__ void pygw_gw0_enk(const double* __restrict__ bands, int nband, const double* __restrict__ core, int ncore,
                             const int* __restrict__ jrk, int nitems, int nb2, double* __restrict__ enk) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)nitems * nb2) return;
  const int it = (int)(t / nb2), ie = (int)(t % nb2);
  enk[t] = ie < ncore ? core[ie] : bands[(size_t)jrk[it] * nband + (ie - ncore)];
}

int selfc_spline_launch(const void* mwm, int nrows, int nb1, int nom, int nb2, const double* x, const double* xp,
                        int nomp, void* out, hipStream_t stream) {
  if (nom > 128 || nom < 3) return PYGW_E_ARG;
  const long n = (long)nrows * nb2;
  if (n > 0x7fffffffL) return PYGW_E_ARG;
  hipLaunchKernelGGL(pygw_selfc_spline_b, dim3((unsigned)n), dim3(128), 0, stream, (const double2*)mwm, nrows, nb1,
                     nom, nb2, x, xp, nomp, (double2*)out);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

// The spline branch's convolution (mwmp given) with the kernel of each
// (e, i w_n, w_l) shared by all rows of a k-point: one block per (k, group
// of SG frequencies), one wave per row ie1 (rows in groups of SR, the block
// looping over the groups), the precise mesh in tiles of SL points.  Per
// (ie2, tile) the block first writes
//   f[g][l] = w_l (e - i w_n) conj(den) / (|den|^2 pi),  den = w_l^2 + (e - i w_n)^2,
// into LDS (the factor of fr_convolution2 without the numerator, fnc.f90:
// 100-115), then each row's lanes add (mwmp[l] - mwm[w_n]) f[g][l] over
// their points.  The terms are those of the reference's sum, each scaled by
// (e - i w_n)/pi before the l sum rather than after it.
namespace sc {
constexpr int SG = 16;   // Matsubara points per block
constexpr int SR = 8;    // rows (waves) per pass
constexpr int SL = 256;  // precise-mesh points per LDS tile
}  // namespace sc

__global__ void __launch_bounds__(512) pygw_selfc_conv_k(const double* __restrict__ enk, int nb1, int nb2,
                                                        const double2* __restrict__ mwm, int nom,
                                                        const double* __restrict__ omega,
                                                        const double2* __restrict__ mwmp, int nomp,
                                                        const double* __restrict__ omp,
                                                        const double* __restrict__ womp,
                                                        double2* __restrict__ out) {
  using namespace sc;
  __shared__ double2 f[SG][SL];
  const int k = blockIdx.x, i0 = blockIdx.y * SG;
  const int ng = nom - i0 < SG ? nom - i0 : SG;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int nt = blockDim.x, nw = nt >> 6;
  const double wmax = omp[nomp - 1];
  for (int r0 = 0; r0 < nb1; r0 += nw) {
    const int ie1 = r0 + wv;
    const bool live = ie1 < nb1;
    const size_t row = (size_t)k * nb1 + (live ? ie1 : 0);
    double2 acc[SG];
#pragma unroll
    for (int g = 0; g < SG; g++) acc[g] = double2{0.0, 0.0};
    for (int ie2 = 0; ie2 < nb2; ie2++) {
      const double e = enk[(size_t)k * nb2 + ie2];
      double2 mn[SG];
#pragma unroll
      for (int g = 0; g < SG; g++)
        mn[g] = (live && g < ng) ? mwm[(((size_t)k * nom + i0 + g) * nb1 + ie1) * nb2 + ie2] : double2{0.0, 0.0};
      const double2* ml_row = mwmp + (row * nb2 + ie2) * nomp;
      for (int l0 = 0; l0 < nomp; l0 += SL) {
        // this lane's points of the tile, loaded first so that their latency
        // overlaps the kernel factors below
        double2 mlv[SL / 64];
#pragma unroll
        for (int q = 0; q < SL / 64; q++) {
          const int l = l0 + lane + 64 * q;
          mlv[q] = (live && l < nomp) ? ml_row[l] : double2{0.0, 0.0};
        }
        __syncthreads();  // the previous tile is consumed
        for (int x = t; x < SG * SL; x += nt) {
          const int g = x / SL, ll = x % SL, l = l0 + ll;
          double2 v = double2{0.0, 0.0};
          if (g < ng && l < nomp) {
            const double wn = omega[i0 + g];
            const double o2 = omp[l] * omp[l];
            const double2 den = double2{o2 + (e * e - wn * wn), -2.0 * e * wn};
            const double d2 = den.x * den.x + den.y * den.y;
            double rr = __builtin_amdgcn_rcp(d2);
            rr = __builtin_fma(rr, __builtin_fma(-d2, rr, 1.0), rr);
            rr = __builtin_fma(rr, __builtin_fma(-d2, rr, 1.0), rr);
            // conj(den) / |den|^2 times w_l (e - i w_n) / pi
            const double s = womp[l] * rr / PI_F;
            const double2 cd = double2{den.x * s, -den.y * s};
            v = cmul(double2{e, -wn}, cd);
          }
          f[g][ll] = v;
        }
        __syncthreads();
        if (live) {
#pragma unroll
          for (int q = 0; q < SL / 64; q++) {
            const int ll = lane + 64 * q;
            if (l0 + ll >= nomp) break;
            const double2 ml = mlv[q];
#pragma unroll
            for (int g = 0; g < SG; g++) {
              const double2 num = double2{ml.x - mn[g].x, ml.y - mn[g].y};
              acc[g] = cadd(acc[g], cmul(num, f[g][ll]));
            }
          }
        }
      }
    }
    if (live) {
      // the atan tail once per (ie2, iom), the ie2 dealt over the row's lanes
      for (int ie2 = lane; ie2 < nb2; ie2 += 64) {
        const double e = enk[(size_t)k * nb2 + ie2];
#pragma unroll
        for (int g = 0; g < SG; g++) {
          if (g < ng) {
            const double2 mn = mwm[(((size_t)k * nom + i0 + g) * nb1 + ie1) * nb2 + ie2];
            if (hypot(mn.x, mn.y) > 1e-16) {
              const double2 eo = double2{e, -omega[i0 + g]};
              const double2 at = catan_(cdiv(double2{wmax, 0.0}, eo));
              acc[g] = cadd(acc[g], cmul(double2{mn.x / PI_F, mn.y / PI_F}, at));
            }
          }
        }
      }
#pragma unroll
      for (int g = 0; g < SG; g++) {
        double x = acc[g].x, y = acc[g].y;
        for (int off = 32; off > 0; off >>= 1) {
          x += __shfl_xor(x, off, 64);
          y += __shfl_xor(y, off, 64);
        }
        if (lane == 0 && g < ng) out[row * nom + i0 + g] = double2{x, y};
      }
    }
  }
}

int selfc_conv_launch(const double* enk, int nrows, int nb1, int nb2, const void* mwm, int nom, const double* omega,
                      const void* mwmp, int nomp, const double* omp, const double* womp, void* out,
                      hipStream_t stream) {
  if (mwmp) {
    // the spline branch: one block per (k, SG frequencies), a wave per row
    const int waves = nb1 < sc::SR ? nb1 : sc::SR;
    hipLaunchKernelGGL(pygw_selfc_conv_k, dim3(nrows / nb1, ceil_div(nom, sc::SG)), dim3(64 * waves), 0, stream, enk,
                       nb1, nb2, (const double2*)mwm, nom, omega, (const double2*)mwmp, nomp, omp, womp,
                       (double2*)out);
    return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
  }
  dim3 grid(nrows, ceil_div(nom, sc::CG));
  hipLaunchKernelGGL(pygw_selfc_conv, grid, dim3(sc::CT), 0, stream, enk, nb1, nb2, (const double2*)mwm, nom, omega,
                     (const double2*)mwmp, nomp, omp, womp, (double2*)out);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

int padecof_launch(const void* zn, const void* gn, int n, int nstate, void* coef, hipStream_t stream) {
  if (n < 1 || n > 64 || nstate < 1) return PYGW_E_ARG;
  hipLaunchKernelGGL(pygw_padecof, dim3(ceil_div(nstate, 64)), dim3(64), 0, stream, (const double2*)zn,
                     (const double2*)gn, n, nstate, (double2*)coef);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

int padeg_launch(const void* x, int m, const void* zn, const void* coef, int n, int nstate, void* gx,
                 hipStream_t stream) {
  if (n < 1 || m < 1 || nstate < 1) return PYGW_E_ARG;
  hipLaunchKernelGGL(pygw_padeg, dim3(ceil_div((long)nstate * m, 256)), dim3(256), 0, stream, (const double2*)x, m,
                     (const double2*)zn, (const double2*)coef, n, nstate, (double2*)gx);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

int pade_qp_launch(const void* sig, int nstate, int nom, const double* omega, int npts, const double* ene,
                   const double* ene0, const double* sigx, const double* vxc, int iop_es, int iop_gw0, double* out,
                   hipStream_t stream) {
  if (npts > 64 || npts < 1 || npts > nom) return PYGW_E_ARG;
  if (iop_es != 0 && (iop_es != -1 || !ene0)) return PYGW_E_ARG;
  hipLaunchKernelGGL(pygw_pade_qp, dim3(ceil_div(nstate, 64)), dim3(64), 0, stream, (const double2*)sig, nstate, nom,
                     omega, npts, ene, ene0, sigx, vxc, iop_es, iop_gw0, out);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

int selfc_qsum_launch(const void* parts, int nq, long n, double alpha, double beta, void* out, hipStream_t stream) {
  long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(pygw_selfc_qsum, dim3((unsigned)blocks), dim3(256), 0, stream, (const double2*)parts, nq, n,
                     alpha, beta, (double2*)out);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

int gw0_enk_launch(const double* bands, int nband, const double* core, int ncore, const int* jrk, int nitems, int nb2,
                   double* enk, hipStream_t stream) {
  const long n = (long)nitems * nb2;
  hipLaunchKernelGGL(pygw_gw0_enk, dim3(ceil_div(n, 256)), dim3(256), 0, stream, bands, nband, core, ncore, jrk,
                     nitems, nb2, enk);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}
