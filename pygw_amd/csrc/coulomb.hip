// The v^(1/2) basis of the Coulomb matrix and its application to the
// product-basis matrix elements, on gfx950:
//
//   Coul_setev (pygw.py:2440, 2505-2545): eigendecomposition of the
//   Hermitian Coulomb matrix Vmat (scipy's eigh, lower triangle), eigenvalues
//   above evtol kept (at q = 0 the eigenvector with the largest overlap with
//   the constant function, pygw.py:2447-2454, is dropped or always kept by
//   iop_coul_x), ordered by decreasing eigenvalue, barcvm[:, m] = v_j sqrt|ev_j|;
//
//   calc_minm (pygw.py:2666, 2684): minm = mmat conj(barcvm), the last step
//   of the matrix elements M the polarization contracts.
//
// The eigensolver is a block cyclic Jacobi method for complex Hermitian
// matrices.  The (padded) index range is cut into blocks of 16; a round of
// the round-robin ordering pairs the blocks, and each 64 x 64 pivot
// submatrix gets one inner cyclic Jacobi sweep in LDS, accumulating its
// unitary U (one workgroup per pair).  One launch then applies all of a
// round's U's: A' = U_u^H A U_v tile by tile (two FP64 MFMA products per
// 32 x 32 tile) and V' = V U_v, both in place; tiles whose U's are all the
// identity are skipped.
// A scalar rotation is applied when its off-diagonal element is above tol
// sqrt(|a_pp a_qq|) and above a floor of 4e-15 ||A||_F (the rounding level
// the tile products leave); the sweeps stop when no pair is above that.  The
// matrix is first scaled by a power of two to ||A||_F in [0.5, 1).
// The GEMM minm = mmat conj(barcvm) runs on v_mfma_f64_16x16x4f64 with four
// real products per complex product.
#include "common.h"

typedef double d4 __attribute__((ext_vector_type(4)));

namespace cl {
constexpr int BS = 16;            // indices per block
constexpr int PB2 = 2 * BS;       // pivot submatrix side
constexpr int NS = BS * BS;       // threads of the solve (one per 2 x 2 block)
constexpr int TW = PB2 / 2;       // update tile side per wave (2 x 2 waves)
constexpr int NF = TW / 16;       // 16 x 16 MFMA tiles per wave side
constexpr double kTol = 1e-14;    // relative rotation threshold
constexpr double kFloor = 4e-15;  // absolute threshold, times ||A||_F
constexpr int kMaxSweeps = 60;

// pair k of round r of the round-robin ordering of n (even) indices
__device__ __forceinline__ void pair_of(int k, int r, int n, int& p, int& q) {
  const int m = n - 1;
  int a, b;
  if (k == 0) {
    a = r;
    b = m;
  } else {
    a = r + k;
    if (a >= m) a -= m;
    b = r - k;
    if (b < 0) b += m;
  }
  p = min(a, b);
  q = max(a, b);
}

struct Rot {
  double cs, sn;  // cos, sin of the real rotation
  double er, ei;  // e = exp(i phi), phi = arg A[p][q]
  double t, ac;   // tan and |A[p][q]| (for the exact diagonal)
  int p, q;
};

// 1/sqrt(x) and 1/x for normal x > 0: the hardware approximations and two
// Newton steps each (quadratic convergence to full double precision)
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = y * fma(-h * y, y, 1.5);
  y = y * fma(-h * y, y, 1.5);
  return y;
}
__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = fma(y, fma(-x, y, 1.0), y);
  y = fma(y, fma(-x, y, 1.0), y);
  return y;
}

// whether the off-diagonal element c of (a, b) is above the thresholds
// (compared as squares: the matrix is scaled to ||A||_F ~ 1)
__device__ __forceinline__ bool above(double a, double b, double ac2, double thr) {
  return ac2 > thr * thr && ac2 > kTol * kTol * fabs(a * b);
}

// J = [[cs, sn], [-sn conj(e), cs conj(e)]] makes J^H [[a, c], [conj c, b]] J
// diagonal; ac = 0 marks the identity (c below the thresholds)
__device__ __forceinline__ Rot make_rot(double a, double b, double2 c, int p, int q, double thr) {
  Rot R;
  R.p = p;
  R.q = q;
  const double ac2 = c.x * c.x + c.y * c.y;
  if (!above(a, b, ac2, thr)) {
    R.cs = 1.0;
    R.sn = 0.0;
    R.er = 1.0;
    R.ei = 0.0;
    R.t = 0.0;
    R.ac = 0.0;
    return R;
  }
  const double ri = rsqrt_nr(ac2);  // 1/|c|
  const double th = 0.5 * (b - a) * ri, at = fabs(th);
  double t = th == 0.0 ? 1.0 : (at > 1e150 ? 0.5 / at : rcp_nr(at + sqrt(fma(at, at, 1.0))));
  t = copysign(t, th == 0.0 ? 1.0 : th);
  R.cs = rsqrt_nr(fma(t, t, 1.0));
  R.sn = t * R.cs;
  R.er = c.x * ri;
  R.ei = c.y * ri;
  R.t = t;
  R.ac = ac2 * ri;
  return R;
}

__device__ __forceinline__ double2 cmul(double2 x, double2 y) {
  return double2{x.x * y.x - x.y * y.y, x.x * y.y + x.y * y.x};
}
__device__ __forceinline__ double2 cadd(double2 x, double2 y) { return double2{x.x + y.x, x.y + y.y}; }
__device__ __forceinline__ double2 cscale(double2 x, double s) { return double2{x.x * s, x.y * s}; }
}  // namespace cl

// W = the lower triangle of Vmat (C order [n0][n0]) mirrored, padded to n
// (even) with zeros; V = I; rn[i] = sum_j |W[i][j]|^2
__global__ void __launch_bounds__(256) pygw_jac_prep(const double2* __restrict__ A, int n0, int n,
                                                    double2* __restrict__ W, double2* __restrict__ V,
                                                    double* __restrict__ rn) {
  const int i = blockIdx.x;
  double s = 0.0;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    double2 x{0.0, 0.0};
    if (i < n0 && j < n0) {
      if (i > j)
        x = A[(size_t)i * n0 + j];
      else if (i < j) {
        const double2 y = A[(size_t)j * n0 + i];
        x = double2{y.x, -y.y};
      } else
        x = double2{A[(size_t)i * n0 + i].x, 0.0};
    }
    W[(size_t)i * n + j] = x;
    V[(size_t)i * n + j] = double2{i == j ? 1.0 : 0.0, 0.0};
    s += x.x * x.x + x.y * x.y;
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) rn[i] = red[0];
}

// s = 2^-e with ||W||_F = f 2^e, f in [0.5, 1) (row sums added in order):
// thr[0] = floor * ||W||_F s (the threshold of the scaled matrix),
// thr[1] = s, thr[2] = 1/s (0 matrix: s = 1)
__global__ void pygw_jac_norm(const double* __restrict__ rn, int n, double* __restrict__ thr) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int i = 0; i < n; i++) s += rn[i];
  const double f = sqrt(s);
  int e = 0;
  if (f > 0.0) frexp(f, &e);
  thr[0] = cl::kFloor * ldexp(f, -e);
  thr[1] = ldexp(1.0, -e);
  thr[2] = ldexp(1.0, e);
}

// W *= thr[1] (exact: a power of two)
__global__ void __launch_bounds__(256) pygw_jac_scale(double2* __restrict__ W, long nn,
                                                     const double* __restrict__ thr) {
  const double s = thr[1];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nn; i += (long)gridDim.x * blockDim.x) {
    const double2 x = W[i];
    W[i] = double2{x.x * s, x.y * s};
  }
}

// one inner cyclic sweep on the 2BS x 2BS pivot submatrix of pair blockIdx.x
// of round r (blocks I < J: local index l < BS is I*BS + l, else J*BS + l -
// BS); U (the product of the sweep's rotations) to Ubuf[blockIdx.x],
// live[blockIdx.x] = whether any rotation was applied (else U = 1).  One
// thread per 2 x 2 block of the submatrix; blocks whose row and column
// rotations are both the identity are left as they are.
__global__ void __launch_bounds__(cl::NS) pygw_bjac_solve(const double2* __restrict__ A, int n, int r,
                                                       const double* __restrict__ thr, double2* __restrict__ Ubuf,
                                                       int* __restrict__ live) {
  using namespace cl;
  __shared__ double2 S[PB2][PB2 + 1];
  __shared__ double2 U[PB2][PB2 + 1];
  __shared__ Rot rot[BS];
  __shared__ int any;
  const int tid = threadIdx.x;
  int I, J;
  pair_of(blockIdx.x, r, n / BS, I, J);
  auto G = [&](int l) { return l < BS ? I * BS + l : J * BS + (l - BS); };
  for (int e = tid; e < PB2 * PB2; e += NS) {
    const int g = e / PB2, d = e % PB2;
    S[g][d] = A[(size_t)G(g) * n + G(d)];
    U[g][d] = double2{g == d ? 1.0 : 0.0, 0.0};
  }
  if (tid == 0) any = 0;
  const double th = thr[0];
  __syncthreads();
  for (int ir = 0; ir < PB2 - 1; ir++) {
    if (tid < BS) {
      int p, q;
      pair_of(tid, ir, PB2, p, q);
      rot[tid] = make_rot(S[p][p].x, S[q][q].x, S[p][q], p, q, th);
      if (rot[tid].ac != 0.0) any = 1;  // ac = 0 marks the identity
    }
    __syncthreads();
    {
      const int u = tid / BS, v = tid % BS;
      const Rot Ru = rot[u], Rv = rot[v];
      if (Ru.ac == 0.0 && Rv.ac == 0.0) goto next;
      const double2 eb{Rv.er, -Rv.ei};
      const double2 j10 = cscale(eb, -Rv.sn), j11 = cscale(eb, Rv.cs);
      {
        const double2 x0 = U[Ru.p][Rv.p], x1 = U[Ru.p][Rv.q], y0 = U[Ru.q][Rv.p], y1 = U[Ru.q][Rv.q];
        U[Ru.p][Rv.p] = cadd(cscale(x0, Rv.cs), cmul(x1, j10));
        U[Ru.p][Rv.q] = cadd(cscale(x0, Rv.sn), cmul(x1, j11));
        U[Ru.q][Rv.p] = cadd(cscale(y0, Rv.cs), cmul(y1, j10));
        U[Ru.q][Rv.q] = cadd(cscale(y0, Rv.sn), cmul(y1, j11));
      }
      if (u == v) {
        const double a = S[Ru.p][Ru.p].x, b = S[Ru.q][Ru.q].x;
        S[Ru.p][Ru.p] = double2{a - Ru.t * Ru.ac, 0.0};
        S[Ru.q][Ru.q] = double2{b + Ru.t * Ru.ac, 0.0};
        S[Ru.p][Ru.q] = double2{0.0, 0.0};
        S[Ru.q][Ru.p] = double2{0.0, 0.0};
        goto next;
      }
      const double2 x00 = S[Ru.p][Rv.p], x01 = S[Ru.p][Rv.q], x10 = S[Ru.q][Rv.p], x11 = S[Ru.q][Rv.q];
      const double2 y00 = cadd(cscale(x00, Rv.cs), cmul(x01, j10)), y01 = cadd(cscale(x00, Rv.sn), cmul(x01, j11));
      const double2 y10 = cadd(cscale(x10, Rv.cs), cmul(x11, j10)), y11 = cadd(cscale(x10, Rv.sn), cmul(x11, j11));
      const double2 eu{Ru.er, Ru.ei};
      const double2 a1 = cscale(eu, -Ru.sn), b1 = cscale(eu, Ru.cs);
      S[Ru.p][Rv.p] = cadd(cscale(y00, Ru.cs), cmul(a1, y10));
      S[Ru.p][Rv.q] = cadd(cscale(y01, Ru.cs), cmul(a1, y11));
      S[Ru.q][Rv.p] = cadd(cscale(y00, Ru.sn), cmul(b1, y10));
      S[Ru.q][Rv.q] = cadd(cscale(y01, Ru.sn), cmul(b1, y11));
    }
  next:
    __syncthreads();
  }
  double2* Uo = Ubuf + (size_t)blockIdx.x * PB2 * PB2;
  for (int e = tid; e < PB2 * PB2; e += NS) Uo[e] = U[e / PB2][e % PB2];
  if (tid == 0) live[blockIdx.x] = any;
}

// apply the round's U's, in place (every element belongs to one tile).
// Workgroups t < nt^2 (nt = n/2BS pairs): the A tile of pairs (t / nt, t % nt),
// A = U_u^H A U_v; the rest: the V tile of rows 2BS (t / nt) .. + 2BS - 1 and the
// columns of pair t % nt, V = V U_v.  Tiles whose U's are all 1 are skipped.
__global__ void __launch_bounds__(256) pygw_bjac_update(double2* __restrict__ A, double2* __restrict__ V, int n,
                                                       int r, const double2* __restrict__ Ubuf,
                                                       const int* __restrict__ live) {
  using namespace cl;
  __shared__ double2 Xs[PB2][PB2 + 1];
  __shared__ double2 Us[PB2][PB2 + 1];
  const int nt = n / PB2, nb = n / BS;
  int t = blockIdx.x;
  const bool isV = t >= nt * nt;
  if (isV) t -= nt * nt;
  const int tu = t / nt, tv = t % nt;
  if (!live[tv] && (isV || !live[tu])) return;
  int Iu, Ju, Iv, Jv;
  pair_of(tu, r, nb, Iu, Ju);
  pair_of(tv, r, nb, Iv, Jv);
  auto Gu = [&](int l) { return isV ? tu * PB2 + l : (l < BS ? Iu * BS + l : Ju * BS + (l - BS)); };
  auto Gv = [&](int l) { return l < BS ? Iv * BS + l : Jv * BS + (l - BS); };
  const double2* X = isV ? V : A;
  const double2* Uv = Ubuf + (size_t)tv * PB2 * PB2;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1, li = lane & 15, lk = lane >> 4;
  for (int e = tid; e < PB2 * PB2; e += 256) {
    const int g = e / PB2, d = e % PB2;
    Xs[g][d] = X[(size_t)Gu(g) * n + Gv(d)];
    Us[g][d] = Uv[e];
  }
  __syncthreads();
  d4 tR[NF][NF], tI[NF][NF];
#pragma unroll
  for (int a = 0; a < NF; a++)
#pragma unroll
    for (int b = 0; b < NF; b++) {
      tR[a][b] = d4{0, 0, 0, 0};
      tI[a][b] = d4{0, 0, 0, 0};
    }
  // T = X U_v: Re = xr ur - xi ui, Im = xr ui + xi ur
  for (int k0 = 0; k0 < PB2; k0 += 4) {
    const int kk = k0 + lk;
    double xr[NF], xi[NF], ur[NF], ui[NF];
#pragma unroll
    for (int s = 0; s < NF; s++) {
      const double2 x = Xs[TW * wr + 16 * s + li][kk];
      const double2 u = Us[kk][TW * wc + 16 * s + li];
      xr[s] = x.x;
      xi[s] = x.y;
      ur[s] = u.x;
      ui[s] = u.y;
    }
#pragma unroll
    for (int a = 0; a < NF; a++)
#pragma unroll
      for (int b = 0; b < NF; b++) {
        tR[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xr[a], ur[b], tR[a][b], 0, 0, 0);
        tR[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-xi[a], ui[b], tR[a][b], 0, 0, 0);
        tI[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xr[a], ui[b], tI[a][b], 0, 0, 0);
        tI[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xi[a], ur[b], tI[a][b], 0, 0, 0);
      }
  }
  if (isV) {
#pragma unroll
    for (int a = 0; a < NF; a++)
#pragma unroll
      for (int b = 0; b < NF; b++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int row = TW * wr + 16 * a + 4 * v + lk, col = TW * wc + 16 * b + li;
          V[(size_t)Gu(row) * n + Gv(col)] = double2{tR[a][b][v], tI[a][b][v]};
        }
    return;
  }
  __syncthreads();  // every wave is done with Xs and Us
  const double2* Uu = Ubuf + (size_t)tu * PB2 * PB2;
#pragma unroll
  for (int a = 0; a < NF; a++)
#pragma unroll
    for (int b = 0; b < NF; b++)
#pragma unroll
      for (int v = 0; v < 4; v++)
        Xs[TW * wr + 16 * a + 4 * v + lk][TW * wc + 16 * b + li] = double2{tR[a][b][v], tI[a][b][v]};
  for (int e = tid; e < PB2 * PB2; e += 256) Us[e / PB2][e % PB2] = Uu[e];
  __syncthreads();
  // Y = U_u^H T: Re = ur tr + ui ti, Im = ur ti - ui tr (u = U_u[gamma][alpha])
#pragma unroll
  for (int a = 0; a < NF; a++)
#pragma unroll
    for (int b = 0; b < NF; b++) {
      tR[a][b] = d4{0, 0, 0, 0};
      tI[a][b] = d4{0, 0, 0, 0};
    }
  for (int k0 = 0; k0 < PB2; k0 += 4) {
    const int kk = k0 + lk;
    double ur[NF], ui[NF], xr[NF], xi[NF];
#pragma unroll
    for (int s = 0; s < NF; s++) {
      const double2 u = Us[kk][TW * wr + 16 * s + li];
      const double2 x = Xs[kk][TW * wc + 16 * s + li];
      ur[s] = u.x;
      ui[s] = u.y;
      xr[s] = x.x;
      xi[s] = x.y;
    }
#pragma unroll
    for (int a = 0; a < NF; a++)
#pragma unroll
      for (int b = 0; b < NF; b++) {
        tR[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ur[a], xr[b], tR[a][b], 0, 0, 0);
        tR[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ui[a], xi[b], tR[a][b], 0, 0, 0);
        tI[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ur[a], xi[b], tI[a][b], 0, 0, 0);
        tI[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ui[a], xr[b], tI[a][b], 0, 0, 0);
      }
  }
#pragma unroll
  for (int a = 0; a < NF; a++)
#pragma unroll
    for (int b = 0; b < NF; b++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = TW * wr + 16 * a + 4 * v + lk, col = TW * wc + 16 * b + li;
        A[(size_t)Gu(row) * n + Gv(col)] = double2{tR[a][b][v], tI[a][b][v]};
      }
}

// count of pairs p < q whose rotation would be applied
__global__ void __launch_bounds__(256) pygw_jac_check(const double2* __restrict__ A, int n,
                                                     const double* __restrict__ thr, unsigned* __restrict__ cnt) {
  const int p = blockIdx.x;
  const double a = A[(size_t)p * n + p].x, th = thr[0];
  unsigned c = 0;
  for (int q = p + 1 + threadIdx.x; q < n; q += blockDim.x) {
    const double2 x = A[(size_t)p * n + q];
    c += cl::above(a, A[(size_t)q * n + q].x, x.x * x.x + x.y * x.y, th);
  }
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(cnt, c);
}

// ascending order of the n0 eigenvalues (ties by index): ev[pos], perm[pos]
__global__ void __launch_bounds__(256) pygw_jac_rank(const double2* __restrict__ A, int n0, int n,
                                                    const double* __restrict__ thr, double* __restrict__ ev,
                                                    int* __restrict__ perm) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n0) return;
  const double x = A[(size_t)j * n + j].x;
  int pos = 0;
  for (int i = 0; i < n0; i++) {
    const double y = A[(size_t)i * n + i].x;
    pos += y < x || (y == x && i < j);
  }
  ev[pos] = x * thr[2];
  perm[pos] = j;
}

// ovl[pos] = | sum_i wi0[i] conj(V[i][perm[pos]]) |  (pygw.py:2450)
__global__ void __launch_bounds__(256) pygw_jac_overlap(const double2* __restrict__ V, int n0, int n,
                                                       const int* __restrict__ perm, const double2* __restrict__ wi0,
                                                       double* __restrict__ ovl) {
  const int pos = blockIdx.x, j = perm[pos];
  double sr = 0.0, si = 0.0;
  for (int i = threadIdx.x; i < n0; i += blockDim.x) {
    const double2 w = wi0[i], v = V[(size_t)i * n + j];
    sr += w.x * v.x + w.y * v.y;
    si += w.y * v.x - w.x * v.y;
  }
  __shared__ double rr[256], ri[256];
  rr[threadIdx.x] = sr;
  ri[threadIdx.x] = si;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      rr[threadIdx.x] += rr[threadIdx.x + o];
      ri[threadIdx.x] += ri[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) ovl[pos] = hypot(rr[0], ri[0]);
}

// kept set and its order (pygw.py:2525-2541), one block of 1024 threads:
// kept = ev > evtol, at q = 0 the largest-overlap eigenvector dropped
// (iop_coul_x = 0) or kept (otherwise); im = matsize - 1 - (kept below).
// colsrc[im] = eigenvector column, barcev[im] = eigenvalue; info = {matsize, immax}.
__global__ void __launch_bounds__(1024) pygw_jac_select(const double* __restrict__ ev, const int* __restrict__ perm,
                                                       int n0, double evtol, const double* __restrict__ ovl,
                                                       int iop_coul_x, int* __restrict__ colsrc,
                                                       double* __restrict__ barcev, int* __restrict__ info) {
  __shared__ int part[1024];
  __shared__ double bo[1024];
  __shared__ int bi[1024];
  const int t = threadIdx.x, T = blockDim.x;
  const int per = (n0 + T - 1) / T, s0 = min(n0, t * per), s1 = min(n0, s0 + per);
  // argmax of the overlap (first of equal maxima)
  int immax = -1;
  if (ovl) {
    double best = -1.0;
    int bj = -1;
    for (int i = s0; i < s1; i++)
      if (ovl[i] > best) {
        best = ovl[i];
        bj = i;
      }
    bo[t] = best;
    bi[t] = bj;
    __syncthreads();
    for (int o = T / 2; o > 0; o >>= 1) {
      if (t < o) {
        const double x = bo[t + o];
        const int y = bi[t + o];
        if (y >= 0 && (x > bo[t] || (x == bo[t] && (bi[t] < 0 || y < bi[t])))) {
          bo[t] = x;
          bi[t] = y;
        }
      }
      __syncthreads();
    }
    immax = bi[0];
  }
  auto kept = [&](int i) { return i == immax ? iop_coul_x != 0 : ev[i] > evtol; };
  int c = 0;
  for (int i = s0; i < s1; i++) c += kept(i);
  part[t] = c;
  __syncthreads();
  if (t == 0) {
    int acc = 0;
    for (int i = 0; i < T; i++) {
      const int x = part[i];
      part[i] = acc;
      acc += x;
    }
    info[0] = acc;
    info[1] = immax;
  }
  __syncthreads();
  const int ms = info[0];
  int below = part[t];
  for (int i = s0; i < s1; i++)
    if (kept(i)) {
      const int im = ms - 1 - below;
      colsrc[im] = perm[i];
      barcev[im] = ev[i];
      below++;
    }
}

// barcvm[i + m ld] = V[i][colsrc[m]] sqrt|barcev[m]| for m < matsize; vecs
// (optional, Fortran [n0, n0]) = eigenvectors in ascending order
__global__ void __launch_bounds__(256) pygw_jac_vectors(const double2* __restrict__ V, int n0, int n,
                                                       const int* __restrict__ colsrc,
                                                       const double* __restrict__ barcev,
                                                       const int* __restrict__ info, double2* __restrict__ barcvm,
                                                       const int* __restrict__ perm, double2* __restrict__ vecs) {
  const int m = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n0) return;
  if (vecs) vecs[(size_t)m * n0 + i] = V[(size_t)i * n + perm[m]];
  if (m < info[0]) {
    const double s = sqrt(fabs(barcev[m]));
    const double2 v = V[(size_t)i * n + colsrc[m]];
    barcvm[(size_t)m * n0 + i] = double2{v.x * s, v.y * s};
  }
}

namespace cl {
// the index range padded to whole pivot submatrices (zero rows and columns)
inline int npad(int n0) { return (n0 + PB2 - 1) / PB2 * PB2; }
// workspace: W[n][n], V[n][n], U[n/2BS][2BS][2BS] complex; rn[n], thr,
// ev[n], ovl[n] double; perm[n], colsrc[n], live[n/2BS], info[2] int, counter
inline size_t ws_bytes(int n0) {
  const size_t n = (size_t)npad(n0);
  return 2 * n * n * 16 + n * PB2 * 16 + (3 * n + 8) * 8 + (3 * n + 8) * 4 + 512;
}
}  // namespace cl

size_t coul_setev_ws_bytes(int mbsize) { return cl::ws_bytes(mbsize); }

int coul_setev_launch(const void* vmat, int n0, double evtol, const void* wi0, int iop_coul_x, double* ev_out,
                      double* barcev, void* barcvm, void* vecs, int* matsize, int* immax, int* sweeps, void* ws,
                      size_t ws_size, hipStream_t stream) {
  using namespace cl;
  if (ws_size < ws_bytes(n0)) return PYGW_E_ARG;
  const int n = npad(n0);
  char* p = (char*)ws;
  double2* W0 = (double2*)p;
  p += (size_t)n * n * 16;
  double2* V = (double2*)p;
  p += (size_t)n * n * 16;
  double2* Ub = (double2*)p;
  p += (size_t)n * PB2 * 16;
  double* rn = (double*)p;
  p += (size_t)n * 8;
  double* thr = (double*)p;
  p += 8 * 8;
  double* ev = (double*)p;
  p += (size_t)n * 8;
  double* ovl = (double*)p;
  p += (size_t)n * 8;
  int* perm = (int*)p;
  p += (size_t)n * 4;
  int* colsrc = (int*)p;
  p += (size_t)n * 4;
  int* live = (int*)p;
  p += (size_t)n * 4;
  int* info = (int*)p;
  p += 8 * 4;
  unsigned* cnt = (unsigned*)(((uintptr_t)p + 255) & ~(uintptr_t)255);

  hipLaunchKernelGGL(pygw_jac_prep, dim3(n), dim3(256), 0, stream, (const double2*)vmat, n0, n, W0, V, rn);
  hipLaunchKernelGGL(pygw_jac_norm, dim3(1), dim3(64), 0, stream, rn, n, thr);
  hipLaunchKernelGGL(pygw_jac_scale, dim3(ceil_div((long)n * n, 1024)), dim3(256), 0, stream, W0, (long)n * n, thr);
  HIP_TRY(hipGetLastError());
  double2* src = W0;
  const int nt = n / PB2, nb = n / BS;
  int sw = 0;
  bool done = false;
  unsigned h = 0;
  // a sweep only starts when some pair is above the threshold
  for (; sw <= kMaxSweeps; sw++) {
    HIP_TRY(hipMemsetAsync(cnt, 0, sizeof(unsigned), stream));
    hipLaunchKernelGGL(pygw_jac_check, dim3(n), dim3(256), 0, stream, src, n, thr, cnt);
    HIP_TRY(hipMemcpyAsync(&h, cnt, sizeof(unsigned), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (h == 0) {
      done = true;
      break;
    }
    if (sw == kMaxSweeps) break;
    for (int r = 0; r < nb - 1; r++) {
      hipLaunchKernelGGL(pygw_bjac_solve, dim3(nt), dim3(NS), 0, stream, src, n, r, thr, Ub, live);
      hipLaunchKernelGGL(pygw_bjac_update, dim3(2 * nt * nt), dim3(256), 0, stream, src, V, n, r, Ub, live);
    }
    HIP_TRY(hipGetLastError());
  }
  if (sweeps) *sweeps = sw;
  if (!done) return PYGW_E_NOCONV;
  hipLaunchKernelGGL(pygw_jac_rank, dim3(ceil_div(n0, 256)), dim3(256), 0, stream, src, n0, n, thr, ev, perm);
  if (wi0)
    hipLaunchKernelGGL(pygw_jac_overlap, dim3(n0), dim3(256), 0, stream, V, n0, n, perm, (const double2*)wi0, ovl);
  hipLaunchKernelGGL(pygw_jac_select, dim3(1), dim3(1024), 0, stream, ev, perm, n0, evtol, wi0 ? ovl : nullptr,
                     iop_coul_x, colsrc, barcev, info);
  hipLaunchKernelGGL(pygw_jac_vectors, dim3(ceil_div(n0, 256), n0), dim3(256), 0, stream, V, n0, n, colsrc, barcev,
                     info, (double2*)barcvm, perm, (double2*)vecs);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(ev_out, ev, (size_t)n0 * 8, hipMemcpyDeviceToDevice, stream));
  int hinfo[2];
  HIP_TRY(hipMemcpyAsync(hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  *matsize = hinfo[0];
  if (immax) *immax = hinfo[1];
  return PYGW_OK;
}

// ---------------------------------------------------------------------------
// minm[r][m] = sum_i mmat[r][i] conj(barcvm[i][m]), i < nin, m < matsize
// (pygw.py:2666; nin = loctmatsize for minc, pygw.py:2684).  mmat row-major
// with leading dimension ldm, barcvm Fortran with leading dimension ldb,
// minm row-major [nrows][matsize], complex128 or rounded to complex64.
// One block: 64 rows x 64 columns, four waves of 32 x 32 (2 x 2 MFMA tiles).
// Items (row strip, column tile) in row-strip order are dealt to the 8 XCDs
// in contiguous runs (block b takes item (b & 7) per + (b >> 3)), so the
// column tiles of a strip run together on one XCD and read the strip from
// its L2.
// ---------------------------------------------------------------------------
namespace vh {
constexpr int TR = 64, TC = 64, KC = 16;
}

// (at least 4 waves per SIMD: the complex64 variant otherwise takes 92 VGPRs
// and runs 3)
template <typename OT>
__global__ void __launch_bounds__(256, 4) pygw_minm_vhalf(const double2* __restrict__ A, long nrows, int nin, long ldm,
                                                      const double2* __restrict__ Bm, long ldb, int matsize,
                                                      long per, OT* __restrict__ C) {
  using namespace vh;
  __shared__ double2 As[TR][KC + 1];
  __shared__ double2 Bs[KC][TC + 1];
  const int nct = (matsize + TC - 1) / TC;
  const long item = (long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (item >= ((nrows + TR - 1) / TR) * nct) return;
  const long r0 = item / nct * TR;
  const int c0 = (int)(item % nct) * TC;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int li = lane & 15, lk = lane >> 4;
  d4 aR[2][2], aI[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      aR[a][b] = d4{0, 0, 0, 0};
      aI[a][b] = d4{0, 0, 0, 0};
    }
  for (int k0 = 0; k0 < nin; k0 += KC) {
    for (int e = tid; e < TR * KC; e += 256) {
      const int r = e / KC, c = e % KC;
      const long row = r0 + r;
      const int k = k0 + c;
      As[r][c] = (row < nrows && k < nin) ? A[row * ldm + k] : double2{0.0, 0.0};
    }
    for (int e = tid; e < KC * TC; e += 256) {
      const int c = e % KC, j = e / KC;
      const int k = k0 + c, m = c0 + j;
      Bs[c][j] = (k < nin && m < matsize) ? Bm[(long)m * ldb + k] : double2{0.0, 0.0};
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
      // a conj(b): Re = ar br + ai bi, Im = ai br - ar bi
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
          aR[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[b], aR[a][b], 0, 0, 0);
          aR[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], bi[b], aR[a][b], 0, 0, 0);
          aI[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], br[b], aI[a][b], 0, 0, 0);
          aI[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ar[a], bi[b], aI[a][b], 0, 0, 0);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const long row = r0 + 32 * wr + 16 * a + 4 * v + lk;
      if (row >= nrows) continue;
#pragma unroll
      for (int b = 0; b < 2; b++) {
        const int m = c0 + 32 * wc + 16 * b + li;
        if (m < matsize) {
          OT o;
          o.x = aR[a][b][v];
          o.y = aI[a][b][v];
          C[row * matsize + m] = o;
        }
      }
    }
}

int minm_vhalf_launch(const void* mmat, long nrows, int nin, long ldm, const void* barcvm, long ldb, int matsize,
                      void* minm, int out_c64, hipStream_t stream) {
  using namespace vh;
  const long items = (nrows + TR - 1) / TR * ((matsize + TC - 1) / TC);
  const long per = (items + 7) / 8;
  if (8 * per > 0x7fffffffL) return PYGW_E_ARG;
  dim3 grid((unsigned)(8 * per));
  if (out_c64)
    hipLaunchKernelGGL(pygw_minm_vhalf<float2>, grid, dim3(256), 0, stream, (const double2*)mmat, nrows, nin, ldm,
                       (const double2*)barcvm, ldb, matsize, per, (float2*)minm);
  else
    hipLaunchKernelGGL(pygw_minm_vhalf<double2>, grid, dim3(256), 0, stream, (const double2*)mmat, nrows, nin, ldm,
                       (const double2*)barcvm, ldb, matsize, per, (double2*)minm);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}
