// Compression of the frequency dependence of the tetrahedron weights
// (pygw_omega_basis / pygw_omega_project / pygw_omega_expand,
// include/pygw_hip.h).
//
// The weights kcw[k, w, p] of one q (pygw.py:3141-3160, layout of
// for_tetrahedra.f90:390-464 transposed to k-major) are, as functions of the
// Matsubara frequency w, combinations of a few smooth profiles: the singular
// values of the (N_k N_p) x n_om matrix of weights fall by seven to eight
// orders of magnitude within about ten terms.  With U the eigenvectors of the
// Gram matrix G = sum_{k,p} kcw[k,:,p] kcw[k,:,p]^T (decreasing eigenvalues)
// and U_R its first R columns, kcw ~ (kcw U_R) U_R^T, and the contraction of
// calc_eps (pygw.py:3161-3200)
//   chi0[I,J,w] = sum_{k,p} kcw[k,w,p] M[k,p,I] conj(M[k,p,J])
// becomes R contractions with the projected weights
//   c[k,r,p] = sum_w U[w,r] kcw[k,w,p]                     (pygw_omega_project)
// into X_r[I,J], followed by the expansion
//   chi0[I,J,w] = sum_r U[w,r] X_r[I,J]                    (pygw_omega_expand).
// This is the idea of the intermediate and discrete Lehmann representations
// of imaginary-frequency functions, with the basis taken from the weights
// themselves.  The truncation changes chi0 by the weights' discarded
// components only; U is orthonormal, so the projection neither amplifies the
// weights nor mixes frequencies beyond what U_R spans.
//
// Balanced basis: the frequencies are first scaled to equal weight (each
// column of the weight matrix to unit norm, D = diag of the column norms),
// U the eigenvectors of that Gram matrix, and then the projection uses
// D^-1 U and the expansion D U.  The truncation and the split-precision
// rounding then cost every frequency the same relative error, also where
// the weights are small (high frequencies of a wide log mesh, where an
// unbalanced basis makes chi0 a cancelling sum of larger terms).
//
// Gram matrix: FP64, each block sums its column tiles into its own partial
// matrix and pygw_wgram_sum adds the partials in block order (deterministic).
// Eigen-decomposition: cyclic Jacobi in one block with G and the accumulated
// rotations in LDS (n_om <= 96), the n/2 disjoint rotations of a round-robin
// step applied together.
#include "common.h"
#include "../../include/pygw_hip.h"

namespace ob {
constexpr int kMaxOm = PYGW_OMEGA_MAX;  // frequencies
constexpr int TP = 64;                  // pair columns per Gram tile
constexpr int kGramBlocks = 512;        // partial Gram matrices
constexpr int RC = 32;                  // basis columns per pass of project / expand
}  // namespace ob

// part[b][i][j] = sum over the column tiles t = b, b + gridDim.x, ... of
// sum_p kcw[k,i,p] kcw[k,j,p] (tile t = k-point k, pairs 64 (t % ntp) ..).
// A thread owns up to three 4 x 4 blocks of the matrix.
__global__ void __launch_bounds__(256) pygw_wgram(const double* __restrict__ kcw, int nk, int nom, int np,
                                                  double* __restrict__ part) {
  using namespace ob;
  __shared__ double T[kMaxOm][TP + 1];
  const int t = threadIdx.x;
  const int ntp = (np + TP - 1) / TP;
  const long ntiles = (long)nk * ntp;
  const int nb4 = (nom + 3) / 4, nblk = nb4 * nb4, rows = 4 * nb4;
  double acc[3][4][4];
#pragma unroll
  for (int s = 0; s < 3; s++)
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int v = 0; v < 4; v++) acc[s][u][v] = 0.0;
  for (long tl = blockIdx.x; tl < ntiles; tl += gridDim.x) {
    const int k = (int)(tl / ntp), p0 = (int)(tl % ntp) * TP;
    __syncthreads();
    for (int e = t; e < rows * TP; e += 256) {
      const int w = e / TP, pp = e % TP, p = p0 + pp;
      T[w][pp] = (w < nom && p < np) ? kcw[((size_t)k * nom + w) * np + p] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 3; s++) {
      const int b = t + 256 * s;
      if (b < nblk) {
        const int i0 = 4 * (b / nb4), j0 = 4 * (b % nb4);
        for (int pp = 0; pp < TP; pp++) {
          double x[4], y[4];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            x[u] = T[i0 + u][pp];
            y[u] = T[j0 + u][pp];
          }
#pragma unroll
          for (int u = 0; u < 4; u++)
#pragma unroll
            for (int v = 0; v < 4; v++) acc[s][u][v] = fma(x[u], y[v], acc[s][u][v]);
        }
      }
    }
  }
  double* o = part + (size_t)blockIdx.x * nom * nom;
#pragma unroll
  for (int s = 0; s < 3; s++) {
    const int b = t + 256 * s;
    if (b >= nblk) continue;
    const int i0 = 4 * (b / nb4), j0 = 4 * (b % nb4);
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int v = 0; v < 4; v++)
        if (i0 + u < nom && j0 + v < nom) o[(size_t)(i0 + u) * nom + j0 + v] = acc[s][u][v];
  }
}

// G[e] = sum_b part[b][e], b = 0 .. nparts-1 in order
__global__ void __launch_bounds__(256) pygw_wgram_sum(const double* __restrict__ part, int nparts, int nn,
                                                      double* __restrict__ G) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= nn) return;
  double s = 0.0;
  for (int b = 0; b < nparts; b++) s += part[(size_t)b * nn + e];
  G[e] = s;
}

// Eigen-decomposition of the symmetric n x n matrix G (n <= kMaxOm) by the
// cyclic Jacobi method (Golub & Van Loan, Matrix Computations, 8.5) with the
// round-robin ordering: in step s of a sweep the pairs are (0, 1 + s mod
// (m-1)) and (1 + (s+i) mod (m-1), 1 + (s-i) mod (m-1)), i = 1 .. m/2-1
// (m = n rounded up to even; index n of an odd n is a dummy), disjoint, so
// their rotations are applied together: all row updates, then all column
// updates of A and of the accumulated rotations V.  Sweeps stop when the
// off-diagonal Frobenius norm is below 1e-15 of the diagonal's.
// With balance, the matrix is first scaled to unit diagonal, G_ij / (d_i d_j)
// with d_i = sqrt(G_ii) (1 where G_ii = 0): for a Gram matrix that is the
// Gram matrix of the columns scaled to unit norm.  For the eigenvalue of
// rank r (decreasing) and its eigenvector V[:, i]: proj[w*n + r] =
// V[w][i] / d_w, expand[w*n + r] = V[w][i] d_w (d = 1 without balance),
// sv[r] = sqrt(max(lambda, 0)); info[0] = sweeps used.
__global__ void __launch_bounds__(1024) pygw_symeig(const double* __restrict__ G, int n, int max_sweeps, int balance,
                                                    double* __restrict__ proj, double* __restrict__ expand,
                                                    double* __restrict__ sv, int* __restrict__ info) {
  using namespace ob;
  __shared__ double A[kMaxOm][kMaxOm + 1];
  __shared__ double V[kMaxOm][kMaxOm + 1];
  __shared__ double rc[kMaxOm / 2], rs[kMaxOm / 2];
  __shared__ int rp[kMaxOm / 2], rq[kMaxOm / 2];
  __shared__ double red[2][16];
  __shared__ double dn[kMaxOm];
  __shared__ int done;
  // block b: the b-th matrix of a batch
  G += (size_t)blockIdx.x * n * n;
  proj += (size_t)blockIdx.x * n * n;
  expand += (size_t)blockIdx.x * n * n;
  sv += (size_t)blockIdx.x * n;
  if (info) info += blockIdx.x;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int m = n + (n & 1), h = m / 2;
  for (int i = t; i < n; i += 1024) {
    const double g = G[(size_t)i * n + i];
    dn[i] = (balance && g > 0.0) ? sqrt(g) : 1.0;
  }
  __syncthreads();
  for (int e = t; e < n * n; e += 1024) {
    const int i = e / n, j = e % n;
    A[i][j] = G[e] / (dn[i] * dn[j]);
    V[i][j] = i == j ? 1.0 : 0.0;
  }
  __syncthreads();
  int sweep = 0;
  for (; sweep < max_sweeps; sweep++) {
    double off = 0.0, dg = 0.0;
    for (int e = t; e < n * n; e += 1024) {
      const int i = e / n, j = e % n;
      const double a = A[i][j];
      if (i == j) dg += a * a;
      else off += a * a;
    }
    for (int o = 32; o > 0; o >>= 1) {
      off += __shfl_xor(off, o, 64);
      dg += __shfl_xor(dg, o, 64);
    }
    if (lane == 0) {
      red[0][wv] = off;
      red[1][wv] = dg;
    }
    __syncthreads();
    if (t == 0) {
      double so = 0.0, sd = 0.0;
      for (int w = 0; w < 16; w++) {
        so += red[0][w];
        sd += red[1][w];
      }
      done = !(so > 1e-30 * sd);
    }
    __syncthreads();
    if (done) break;
    for (int s = 0; s < m - 1; s++) {
      if (t < h) {
        int p, q;
        if (t == 0) {
          p = 0;
          q = 1 + s % (m - 1);
        } else {
          p = 1 + (s + t) % (m - 1);
          q = 1 + (s - t + (m - 1)) % (m - 1);
        }
        if (p > q) {
          const int x = p;
          p = q;
          q = x;
        }
        double c = 1.0, sn = 0.0;
        if (q < n) {
          const double apq = A[p][q];
          if (apq != 0.0) {
            const double tau = (A[q][q] - A[p][p]) / (2.0 * apq);
            const double tt = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + tt * tt);
            sn = tt * c;
          }
        }
        rp[t] = p;
        rq[t] = q;
        rc[t] = c;
        rs[t] = sn;
      }
      __syncthreads();
      for (int e = t; e < h * n; e += 1024) {  // rows p, q of J^T A
        const int i = e / n, j = e % n;
        if (rs[i] == 0.0) continue;
        const int p = rp[i], q = rq[i];
        const double c = rc[i], sn = rs[i], ap = A[p][j], aq = A[q][j];
        A[p][j] = c * ap - sn * aq;
        A[q][j] = sn * ap + c * aq;
      }
      __syncthreads();
      for (int e = t; e < h * n; e += 1024) {  // columns p, q of (J^T A) J and of V J
        const int i = e / n, j = e % n;
        if (rs[i] == 0.0) continue;
        const int p = rp[i], q = rq[i];
        const double c = rc[i], sn = rs[i];
        const double ap = A[j][p], aq = A[j][q];
        A[j][p] = c * ap - sn * aq;
        A[j][q] = sn * ap + c * aq;
        const double vp = V[j][p], vq = V[j][q];
        V[j][p] = c * vp - sn * vq;
        V[j][q] = sn * vp + c * vq;
      }
      __syncthreads();
      if (t < h && rs[t] != 0.0) {  // the rotation annihilates A[p][q]
        A[rp[t]][rq[t]] = 0.0;
        A[rq[t]][rp[t]] = 0.0;
      }
      __syncthreads();
    }
  }
  for (int i = t; i < n; i += 1024) {
    const double li = A[i][i];
    int r = 0;
    for (int j = 0; j < n; j++) {
      const double lj = A[j][j];
      r += (lj > li || (lj == li && j < i)) ? 1 : 0;
    }
    for (int w = 0; w < n; w++) {
      proj[(size_t)w * n + r] = V[w][i] / dn[w];
      expand[(size_t)w * n + r] = V[w][i] * dn[w];
    }
    sv[r] = sqrt(fmax(li, 0.0));
  }
  if (t == 0 && info) info[0] = sweep;
}

// out[k][r][p] = sum_w P(w, r) kcw[k][w][p], r < rank, P(w, r) = proj[w*sw +
// r*sr]; a thread per (k, p)
__global__ void __launch_bounds__(256) pygw_wproject(const double* __restrict__ kcw, int nk, int nom, int np,
                                                     const double* __restrict__ proj, long sw, long sr, int rank,
                                                     double* __restrict__ out) {
  using namespace ob;
  __shared__ double Us[kMaxOm * RC];
  const long ncol = (long)nk * np;
  for (int r0 = 0; r0 < rank; r0 += RC) {
    const int nr = rank - r0 < RC ? rank - r0 : RC;
    __syncthreads();
    for (int e = threadIdx.x; e < nom * RC; e += 256) {
      const int w = e / RC, rr = e % RC;
      Us[e] = rr < nr ? proj[w * sw + (r0 + rr) * sr] : 0.0;
    }
    __syncthreads();
    for (long col = blockIdx.x * 256L + threadIdx.x; col < ncol; col += (long)gridDim.x * 256) {
      const int k = (int)(col / np), p = (int)(col % np);
      double acc[RC];
#pragma unroll
      for (int rr = 0; rr < RC; rr++) acc[rr] = 0.0;
      const double* src = kcw + (size_t)k * nom * np + p;
      for (int w = 0; w < nom; w++) {
        const double x = src[(size_t)w * np];
#pragma unroll
        for (int rr = 0; rr < RC; rr++) acc[rr] = fma(Us[w * RC + rr], x, acc[rr]);
      }
      double* dst = out + ((size_t)k * rank + r0) * np + p;
#pragma unroll
      for (int rr = 0; rr < RC; rr++)
        if (rr < nr) dst[(size_t)rr * np] = acc[rr];
    }
  }
}

// eps[b][w][e] = (beta ? eps[b][w][e] : 0) + sum_r E(w, r) X[b][r][e],
// e < nel, batch b = blockIdx.y, E(w, r) = ex[w*sw + r*sr]; a thread per
// element e (the same e in every pass over r, so the passes after the first
// add to what the first wrote)
__global__ void __launch_bounds__(256) pygw_omega_expand_k(const double2* __restrict__ X, const double* __restrict__ ex,
                                                           long sw, long sr, long nel, int rank, int nom, int beta,
                                                           double2* __restrict__ eps) {
  using namespace ob;
  __shared__ double Us[kMaxOm * RC];
  X += (size_t)blockIdx.y * rank * nel;
  eps += (size_t)blockIdx.y * nom * nel;
  for (int r0 = 0; r0 < rank; r0 += RC) {
    const int nr = rank - r0 < RC ? rank - r0 : RC;
    __syncthreads();
    for (int e = threadIdx.x; e < nom * RC; e += 256) {
      const int w = e / RC, rr = e % RC;
      Us[e] = rr < nr ? ex[w * sw + (r0 + rr) * sr] : 0.0;
    }
    __syncthreads();
    const bool add = beta || r0 > 0;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < nel; e += (long)gridDim.x * 256) {
      double2 x[RC];
#pragma unroll
      for (int rr = 0; rr < RC; rr++) x[rr] = rr < nr ? X[(size_t)(r0 + rr) * nel + e] : double2{0.0, 0.0};
      for (int w = 0; w < nom; w++) {
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int rr = 0; rr < RC; rr++) {
          re = fma(Us[w * RC + rr], x[rr].x, re);
          im = fma(Us[w * RC + rr], x[rr].y, im);
        }
        double2* o = eps + (size_t)w * nel + e;
        if (add) {
          const double2 v = *o;
          *o = double2{v.x + re, v.y + im};
        } else {
          *o = double2{re, im};
        }
      }
    }
  }
}

namespace {
size_t basis_ws(int nom) { return ((size_t)ob::kGramBlocks + 1) * nom * nom * sizeof(double) + 256; }
}  // namespace

extern "C" int pygw_omega_basis_workspace(int n_k, int n_om, int n_pairs, size_t* bytes) {
  if (n_k <= 0 || n_pairs <= 0 || n_om <= 0 || n_om > ob::kMaxOm || !bytes) return PYGW_E_ARG;
  *bytes = basis_ws(n_om);
  return PYGW_OK;
}

extern "C" int pygw_omega_basis_many(int n_q, const double* const* kcw, const int* n_k, const int* n_pairs, int n_om,
                                     int balance, double* proj, double* expand, double* sv, int* info,
                                     void* workspace, size_t ws_bytes, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (n_q <= 0 || n_q > 65535 || !kcw || !n_k || !n_pairs || !proj || !expand || !sv || n_om <= 0 ||
      n_om > ob::kMaxOm || (balance != 0 && balance != 1))
    return PYGW_E_ARG;
  for (int q = 0; q < n_q; q++)
    if (!kcw[q] || n_k[q] <= 0 || n_pairs[q] <= 0) return PYGW_E_ARG;
  const size_t nn = (size_t)n_om * n_om;
  const size_t need = basis_ws(n_om) + (size_t)(n_q - 1) * nn * sizeof(double);
  void* ws = workspace;
  bool own = false;
  if (!ws) {
    HIP_TRY(hipMallocAsync(&ws, need, stream));
    own = true;
  } else if (ws_bytes < need) {
    return PYGW_E_ARG;
  }
  // partial sums (reused q after q, stream-ordered), then the n_q Gram matrices
  double* part = (double*)ws;
  double* G = part + (size_t)ob::kGramBlocks * nn;
  for (int q = 0; q < n_q; q++) {
    const long ntiles = (long)n_k[q] * ((n_pairs[q] + ob::TP - 1) / ob::TP);
    const int nparts = ntiles < ob::kGramBlocks ? (int)ntiles : ob::kGramBlocks;
    hipLaunchKernelGGL(pygw_wgram, dim3(nparts), dim3(256), 0, stream, kcw[q], n_k[q], n_om, n_pairs[q], part);
    hipLaunchKernelGGL(pygw_wgram_sum, dim3(ceil_div((long)nn, 256)), dim3(256), 0, stream, (const double*)part,
                       nparts, (int)nn, G + (size_t)q * nn);
  }
  hipLaunchKernelGGL(pygw_symeig, dim3(n_q), dim3(1024), 0, stream, (const double*)G, n_om, 40, balance, proj, expand,
                     sv, info);
  int rc = hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
  if (own && hipFreeAsync(ws, stream) != hipSuccess && rc == PYGW_OK) rc = PYGW_E_HIP;
  return rc;
}

extern "C" int pygw_omega_basis(const double* kcw, int n_k, int n_om, int n_pairs, int balance, double* proj,
                                double* expand, double* sv, int* info, void* workspace, size_t ws_bytes,
                                void* stream) {
  if (!kcw) return PYGW_E_ARG;
  const double* k1[1] = {kcw};
  return pygw_omega_basis_many(1, k1, &n_k, &n_pairs, n_om, balance, proj, expand, sv, info, workspace, ws_bytes,
                               stream);
}

extern "C" int pygw_omega_basis_many_workspace(int n_q, int n_om, size_t* bytes) {
  if (n_q <= 0 || n_om <= 0 || n_om > ob::kMaxOm || !bytes) return PYGW_E_ARG;
  *bytes = basis_ws(n_om) + (size_t)(n_q - 1) * n_om * n_om * sizeof(double);
  return PYGW_OK;
}

extern "C" int pygw_omega_project(const double* kcw, int n_k, int n_om, int n_pairs, const double* proj, long sw,
                                  long sr, int rank, double* kcw_r, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (!kcw || !proj || !kcw_r || n_k <= 0 || n_pairs <= 0 || n_om <= 0 || n_om > ob::kMaxOm || rank < 1 ||
      rank > ob::kMaxOm || sw < 0 || sr < 0)
    return PYGW_E_ARG;
  const long ncol = (long)n_k * n_pairs;
  int blocks = ceil_div(ncol, 256);
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(pygw_wproject, dim3(blocks), dim3(256), 0, stream, kcw, n_k, n_om, n_pairs, proj, sw, sr, rank,
                     kcw_r);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

extern "C" int pygw_omega_expand_batched(const void* x, const double* ex, long sw, long sr, int n_batch, long nel,
                                         int rank, int n_om, int beta, void* out, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (!x || !ex || !out || n_batch <= 0 || n_batch > 65535 || nel <= 0 || n_om <= 0 || n_om > ob::kMaxOm ||
      rank < 1 || rank > ob::kMaxOm || sw < 0 || sr < 0 || (beta != 0 && beta != 1))
    return PYGW_E_ARG;
  int blocks = ceil_div(nel, 256);
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(pygw_omega_expand_k, dim3(blocks, n_batch), dim3(256), 0, stream, (const double2*)x, ex, sw, sr,
                     nel, rank, n_om, beta, (double2*)out);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

extern "C" int pygw_omega_expand(const void* x, const double* ex, long sw, long sr, int pb, int rank, int n_om,
                                 int beta, void* eps, void* stream) {
  if (pb <= 0) return PYGW_E_ARG;
  return pygw_omega_expand_batched(x, ex, sw, sr, 1, (long)pb * pb, rank, n_om, beta, eps, stream);
}
