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
// The eigensolver is a parallel cyclic Jacobi method for complex Hermitian
// matrices: one launch per round of the round-robin ordering (n/2 disjoint
// index pairs), each thread owning one 2 x 2 block (u, v) of the rotated
// matrix A' = J_u^H A J_v (out of place, ping-pong buffers) and of the
// eigenvectors V' = V J_v (in place).  A rotation is applied when its
// off-diagonal element is above tol sqrt(|a_pp a_qq|) and above a floor
// of 1e-16 ||A||_F; the sweeps stop when no pair is above the threshold.
// The GEMM minm = mmat conj(barcvm) runs on v_mfma_f64_16x16x4f64 with four
// real products per complex product.
#include "common.h"

typedef double d4 __attribute__((ext_vector_type(4)));

namespace cl {
constexpr int TP = 16;          // pairs per tile side in a Jacobi round
constexpr double kTol = 1e-14;  // relative rotation threshold
constexpr double kFloor = 1e-16;  // absolute threshold, times ||A||_F
constexpr int kMaxSweeps = 80;

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

// J = [[cs, sn], [-sn conj(e), cs conj(e)]] makes J^H [[a, c], [conj c, b]] J diagonal
__device__ __forceinline__ Rot make_rot(const double2* __restrict__ A, int n, int p, int q, double thr) {
  Rot R;
  R.p = p;
  R.q = q;
  const double a = A[(size_t)p * n + p].x, b = A[(size_t)q * n + q].x;
  const double2 c = A[(size_t)p * n + q];
  const double ac = hypot(c.x, c.y);
  if (ac <= thr || ac <= kTol * sqrt(fabs(a * b))) {
    R.cs = 1.0;
    R.sn = 0.0;
    R.er = 1.0;
    R.ei = 0.0;
    R.t = 0.0;
    R.ac = 0.0;
    return R;
  }
  const double th = (b - a) / (2.0 * ac);
  const double t = th == 0.0 ? 1.0 : copysign(1.0, th) / (fabs(th) + sqrt(th * th + 1.0));
  R.cs = 1.0 / sqrt(1.0 + t * t);
  R.sn = t * R.cs;
  R.er = c.x / ac;
  R.ei = c.y / ac;
  R.t = t;
  R.ac = ac;
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

// thr[0] = floor * ||W||_F (row sums added in order)
__global__ void pygw_jac_norm(const double* __restrict__ rn, int n, double* __restrict__ thr) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int i = 0; i < n; i++) s += rn[i];
  thr[0] = cl::kFloor * sqrt(s);
}

// one round r: B = J^H A J for the n/2 pairs of the round, V = V J
__global__ void __launch_bounds__(256) pygw_jac_round(const double2* __restrict__ A, double2* __restrict__ B,
                                                     double2* __restrict__ V, int n, int r,
                                                     const double* __restrict__ thr) {
  using namespace cl;
  __shared__ Rot ru[TP], rv[TP];
  const int np = n / 2;
  const int u0 = blockIdx.y * TP, v0 = blockIdx.x * TP;
  const double th = thr[0];
  if (threadIdx.x < 2 * TP) {
    const int k = (threadIdx.x < TP ? u0 : v0) + (threadIdx.x % TP);
    if (k < np) {
      int p, q;
      pair_of(k, r, n, p, q);
      const Rot R = make_rot(A, n, p, q, th);
      if (threadIdx.x < TP)
        ru[threadIdx.x] = R;
      else
        rv[threadIdx.x - TP] = R;
    }
  }
  __syncthreads();
  const int tu = threadIdx.x / TP, tv = threadIdx.x % TP;
  const int u = u0 + tu, v = v0 + tv;
  if (u >= np || v >= np) return;
  const Rot U = ru[tu], W = rv[tv];
  const size_t pp = (size_t)U.p * n + W.p, pq = (size_t)U.p * n + W.q;
  const size_t qp = (size_t)U.q * n + W.p, qq = (size_t)U.q * n + W.q;
  // V J_v on rows p_u, q_u
  {
    const double2 eb{W.er, -W.ei};                       // conj(e_v)
    const double2 j10 = cscale(eb, -W.sn), j11 = cscale(eb, W.cs);
    const double2 x0 = V[pp], x1 = V[pq], y0 = V[qp], y1 = V[qq];
    V[pp] = cadd(cscale(x0, W.cs), cmul(x1, j10));
    V[pq] = cadd(cscale(x0, W.sn), cmul(x1, j11));
    V[qp] = cadd(cscale(y0, W.cs), cmul(y1, j10));
    V[qq] = cadd(cscale(y0, W.sn), cmul(y1, j11));
  }
  if (u == v) {
    B[pp] = double2{A[pp].x - U.t * U.ac, 0.0};
    B[qq] = double2{A[qq].x + U.t * U.ac, 0.0};
    B[pq] = double2{0.0, 0.0};
    B[qp] = double2{0.0, 0.0};
    return;
  }
  const double2 x00 = A[pp], x01 = A[pq], x10 = A[qp], x11 = A[qq];
  // Y = X J_v
  const double2 eb{W.er, -W.ei};
  const double2 j10 = cscale(eb, -W.sn), j11 = cscale(eb, W.cs);
  const double2 y00 = cadd(cscale(x00, W.cs), cmul(x01, j10)), y01 = cadd(cscale(x00, W.sn), cmul(x01, j11));
  const double2 y10 = cadd(cscale(x10, W.cs), cmul(x11, j10)), y11 = cadd(cscale(x10, W.sn), cmul(x11, j11));
  // J_u^H Y: row 0 = cs Y0 - sn e Y1, row 1 = sn Y0 + cs e Y1
  const double2 e{U.er, U.ei};
  const double2 a1 = cscale(e, -U.sn), b1 = cscale(e, U.cs);
  B[pp] = cadd(cscale(y00, U.cs), cmul(a1, y10));
  B[pq] = cadd(cscale(y01, U.cs), cmul(a1, y11));
  B[qp] = cadd(cscale(y00, U.sn), cmul(b1, y10));
  B[qq] = cadd(cscale(y01, U.sn), cmul(b1, y11));
}

// count of pairs p < q whose rotation would be applied
__global__ void __launch_bounds__(256) pygw_jac_check(const double2* __restrict__ A, int n,
                                                     const double* __restrict__ thr, unsigned* __restrict__ cnt) {
  const int p = blockIdx.x;
  const double a = A[(size_t)p * n + p].x, th = thr[0];
  unsigned c = 0;
  for (int q = p + 1 + threadIdx.x; q < n; q += blockDim.x) {
    const double2 x = A[(size_t)p * n + q];
    const double ac = hypot(x.x, x.y);
    c += ac > th && ac > cl::kTol * sqrt(fabs(a * A[(size_t)q * n + q].x));
  }
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(cnt, c);
}

// ascending order of the n0 eigenvalues (ties by index): ev[pos], perm[pos]
__global__ void __launch_bounds__(256) pygw_jac_rank(const double2* __restrict__ A, int n0, int n,
                                                    double* __restrict__ ev, int* __restrict__ perm) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n0) return;
  const double x = A[(size_t)j * n + j].x;
  int pos = 0;
  for (int i = 0; i < n0; i++) {
    const double y = A[(size_t)i * n + i].x;
    pos += y < x || (y == x && i < j);
  }
  ev[pos] = x;
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
inline int npad(int n0) { return n0 + (n0 & 1); }
// workspace: W[2][n][n], V[n][n] complex; rn[n], thr, ev[n], ovl[n] double;
// perm[n], colsrc[n], info[2] int, counter
inline size_t ws_bytes(int n0) {
  const size_t n = (size_t)npad(n0);
  return 3 * n * n * 16 + (3 * n + 8) * 8 + (2 * n + 8) * 4 + 512;
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
  double2* W1 = (double2*)p;
  p += (size_t)n * n * 16;
  double2* V = (double2*)p;
  p += (size_t)n * n * 16;
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
  int* info = (int*)p;
  p += 8 * 4;
  unsigned* cnt = (unsigned*)(((uintptr_t)p + 255) & ~(uintptr_t)255);

  hipLaunchKernelGGL(pygw_jac_prep, dim3(n), dim3(256), 0, stream, (const double2*)vmat, n0, n, W0, V, rn);
  hipLaunchKernelGGL(pygw_jac_norm, dim3(1), dim3(64), 0, stream, rn, n, thr);
  HIP_TRY(hipGetLastError());
  double2* src = W0;
  double2* dst = W1;
  const int tiles = ceil_div(n / 2, TP);
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
    for (int r = 0; r < n - 1; r++) {
      hipLaunchKernelGGL(pygw_jac_round, dim3(tiles, tiles), dim3(256), 0, stream, src, dst, V, n, r, thr);
      double2* t = src;
      src = dst;
      dst = t;
    }
    HIP_TRY(hipGetLastError());
  }
  if (sweeps) *sweeps = sw;
  if (!done) return PYGW_E_NOCONV;
  hipLaunchKernelGGL(pygw_jac_rank, dim3(ceil_div(n0, 256)), dim3(256), 0, stream, src, n0, n, ev, perm);
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
// ---------------------------------------------------------------------------
namespace vh {
constexpr int TR = 64, TC = 64, KC = 16;
}

template <typename OT>
__global__ void __launch_bounds__(256) pygw_minm_vhalf(const double2* __restrict__ A, long nrows, int nin, long ldm,
                                                      const double2* __restrict__ Bm, long ldb, int matsize,
                                                      OT* __restrict__ C) {
  using namespace vh;
  __shared__ double2 As[TR][KC + 1];
  __shared__ double2 Bs[KC][TC + 1];
  const long r0 = (long)blockIdx.x * TR;
  const int c0 = blockIdx.y * TC;
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
  const long gx = (nrows + TR - 1) / TR;
  if (gx > 0x7fffffffL) return PYGW_E_ARG;
  dim3 grid((unsigned)gx, ceil_div(matsize, TC));
  if (out_c64)
    hipLaunchKernelGGL(pygw_minm_vhalf<float2>, grid, dim3(256), 0, stream, (const double2*)mmat, nrows, nin, ldm,
                       (const double2*)barcvm, ldb, matsize, (float2*)minm);
  else
    hipLaunchKernelGGL(pygw_minm_vhalf<double2>, grid, dim3(256), 0, stream, (const double2*)mmat, nrows, nin, ldm,
                       (const double2*)barcvm, ldb, matsize, (double2*)minm);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}
