// Split-precision product of the matrix elements with the v^(1/2) basis
// (minm = la.matmul(mmat, conj(barcvm)), pygw.py:2666, 2684) on gfx950 MFMA
// (v_mfma_f32_16x16x32_f16), for the split-precision polarization, which
// contracts minm in complex64 anyway:
//
//   minm[r][m] = sum_{i < n_in} mmat[r][i] conj(barcvm[i][m])
//
// Both operands are split into FP16 hi/lo fragment planes: the rows of mmat
// scaled per row by 2^(14-e_r), the columns of barcvm per column by
// 2^(14-e_m) (e the exponent of the largest |Re| or |Im|), so that every
// value is below 2^14 and its low half stays in FP16's normal range.  A real
// product takes three MFMA terms (hi hi + hi lo + lo hi) into FP32
// accumulators that restart every 4 chunks of 32 in i and are added into a
// second set of FP32 sums: each sum then takes 8 roundings at mbsize 1000
// where one accumulator over the whole sum took 96, which cuts the
// accumulation error ~3x (FP64 sums would cut it further, but their
// registers spill at two waves per SIMD); with a = mmat, b = barcvm:
//   Re = a_r b_r + a_i b_i,   -Im = a_r b_i + (-a_i) b_r.
// The epilogue unscales exactly (powers of two) and stores complex64 or
// complex128.  barcvm's planes are built once per basis
// (pygw_minm_vhalf_split_basis), the rows' planes per call.
//
// Tiles of 128 rows x 128 columns of minm, 512 threads (8 waves of 64 x 32),
// chunks of 32 in i double-buffered through LDS.  The column tiles of one row
// tile are dealt to one XCD in a row, so the row planes are read from HBM
// once per row tile and the basis planes (8 MB at mbsize 1000) stay in the
// L2 / last-level cache.
#include <algorithm>

#include "common.h"
#include "../../include/pygw_hip.h"

#pragma clang fp contract(off)

typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

namespace vh {
constexpr int TB = 128;  // rows and columns per block
constexpr int KC = 32;   // i per chunk = MFMA K
constexpr int NT = 512;  // threads per block
constexpr int FOLD = 4;  // chunks per accumulator run
constexpr int RSRC_FLAGS = 0x00020000;

__device__ __forceinline__ void split(float x, _Float16& hi, _Float16& lo) {
  asm volatile("" : "+v"(x));  // one f32 rounding of x feeds both halves
  hi = (_Float16)x;
  lo = (_Float16)(x - (float)hi);
}

__device__ __forceinline__ h8 neg(h8 x) {
  u4 u = __builtin_bit_cast(u4, x);
  u ^= u4{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
  return __builtin_bit_cast(h8, u);
}

inline size_t al256(size_t x) { return (x + 255) / 256 * 256; }
inline int npc_of(int n) { return (n + KC - 1) / KC; }
inline int pad_of(int n) { return (n + TB - 1) / TB * TB; }
// basis: column scales, then the planes of npc(mb) chunks x 16 x colp
inline size_t basis_bytes(int mb, int ms) {
  return al256(sizeof(float) * (size_t)ms) + (size_t)npc_of(mb) * 16 * pad_of(ms) * sizeof(h8);
}
// per call: row scales, then the planes of npc(n_in) chunks x 16 x rowp
inline size_t rows_bytes(long nrows, int n_in) {
  return al256(sizeof(float) * (size_t)nrows) + (size_t)npc_of(n_in) * 16 * (size_t)pad_of((int)nrows) * sizeof(h8);
}
}  // namespace vh

// scale = 2^(14-e), e the exponent of the largest |Re| or |Im| of the n
// values at v[k * stride], k < n (1 for a zero vector); one block each
__global__ void __launch_bounds__(256) pygw_vh_scale(const double2* __restrict__ v, long vec_stride, long stride,
                                                    int n, float* __restrict__ scale) {
  const double2* x = v + (size_t)blockIdx.x * vec_stride;
  double m = 0.0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) m = fmax(m, fmax(fabs(x[(size_t)k * stride].x), fabs(x[(size_t)k * stride].y)));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    int e = 0;
    if (m > 0.0) frexp(m, &e);
    scale[blockIdx.x] = (float)ldexp(1.0, 14 - e);
  }
}

// Planes P[((c*4 + g)*4 + q)*vp + t] of the vectors t < nv (padded to vp):
// the 8 FP16 values (q: re_hi, re_lo, im_hi, im_lo) of entries i = 32c + 8g
// + j, j < 8, of vector t scaled by s[t], read at v[t * vec_stride + i];
// entries i >= n and vectors t >= nv are zero.  The rows of mmat and the
// columns of barcvm (Fortran order) are both such vectors.  One block per
// 64 vectors x 64 entries: the entries are read along i (coalesced), scaled
// and rounded to FP32 into LDS, then split and written along t (each plane
// row of the tile is 1 KB contiguous).
__global__ void __launch_bounds__(256) pygw_vh_split(const double2* __restrict__ v, long nv, long vec_stride, int n,
                                                    int vp, int npc, const float* __restrict__ s,
                                                    h8* __restrict__ P) {
  __shared__ float2 tile[64][64 + 1];
  const long t0 = (long)blockIdx.x * 64;
  const int i0 = blockIdx.y * 64;
  const int tid = threadIdx.x;
  for (int e = tid; e < 64 * 64; e += 256) {
    const int ti = e >> 6, ii = e & 63;        // consecutive threads: consecutive i
    const long t = t0 + ti;
    const int i = i0 + ii;
    float2 x = float2{0.f, 0.f};
    if (t < nv && i < n) {
      const double2 a = v[(size_t)t * vec_stride + i];
      const double sc = (double)s[t];
      x = float2{(float)(a.x * sc), (float)(a.y * sc)};
    }
    tile[ti][ii] = x;
  }
  __syncthreads();
  // 64 vectors x 8 groups of 8 entries; consecutive threads: consecutive t
  for (int e = tid; e < 64 * 8; e += 256) {
    const int ti = e & 63, gl = e >> 6;
    const long t = t0 + ti;
    if (t >= vp) continue;
    const int cg = (i0 >> 3) + gl;             // c*4 + g
    if (cg >= npc * 4) continue;
    h8 rh, rl, ih, il;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const float2 x = tile[ti][8 * gl + j];
      _Float16 h, l;
      vh::split(x.x, h, l);
      rh[j] = h;
      rl[j] = l;
      vh::split(x.y, h, l);
      ih[j] = h;
      il[j] = l;
    }
    h8* o = P + (size_t)cg * 4 * vp + t;
    o[0] = rh;
    o[(size_t)vp] = rl;
    o[2 * (size_t)vp] = ih;
    o[3 * (size_t)vp] = il;
  }
}

// MFMA 16x16x32 f16 layouts: A[i][k]: i = l%16, k = 8*(l/16) + j;
// B[k][n]: n = l%16, k = 8*(l/16) + j;  D[i][n]: n = l%16, i = 4*(l/16) + v.
// Block b runs on XCD b % 8 as its (b / 8)-th block; item (b % 8) per + b / 8
// = (row tile, column tile), column tiles innermost.
template <bool C64>
__global__ void __launch_bounds__(512) pygw_vh_mixed(const h8* __restrict__ SA, const h8* __restrict__ SB,
                                                    long nrows, int ms, int rowp, int colp, int npc,
                                                    const float* __restrict__ rs, const float* __restrict__ cs,
                                                    int nct, long items, long per, void* __restrict__ out) {
  using namespace vh;
  __shared__ h8 As[2][16][TB];  // [buf][g*4 + plane][row]
  __shared__ h8 Bs[2][16][TB];  // [buf][g*4 + plane][column]
  const long item = (long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (item >= items) return;     // the whole block leaves before any barrier
  const long rt = item / nct;
  const int ct = (int)(item % nct);
  const long r0 = rt * TB;
  const int c0 = ct * TB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 2, wc = wave & 3;
  const int li = lane & 15, lk = lane >> 4;

  // staging: 4 x 16 B of each slab per thread, planes contiguous in memory
  const unsigned aoffB = (unsigned)((tid >> 7) * rowp + (tid & 127)) * 16u;
  const unsigned boffB = (unsigned)((tid >> 7) * colp + (tid & 127)) * 16u;
  const unsigned aqB = 4u * (unsigned)rowp * 16u, bqB = 4u * (unsigned)colp * 16u;
  h8 pa[4], pbv[4];
  auto fetch = [&](int c) {
    const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(SA + (size_t)c * 16 * rowp + r0), (short)0, (int)(16u * (unsigned)rowp * 16u), RSRC_FLAGS);
    const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(SB + (size_t)c * 16 * colp + c0), (short)0, (int)(16u * (unsigned)colp * 16u), RSRC_FLAGS);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      pa[q] = __builtin_bit_cast(h8, __builtin_amdgcn_raw_buffer_load_b128(ra, aoffB + q * aqB, 0, 0));
      pbv[q] = __builtin_bit_cast(h8, __builtin_amdgcn_raw_buffer_load_b128(rb, boffB + q * bqB, 0, 0));
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int idx = tid + NT * q;
      As[buf][idx >> 7][idx & 127] = pa[q];
      Bs[buf][idx >> 7][idx & 127] = pbv[q];
    }
  };

  f4 accR[4][2], accI[4][2], sumR[4][2], sumI[4][2];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      accR[a][b] = f4{0.f, 0.f, 0.f, 0.f};
      accI[a][b] = f4{0.f, 0.f, 0.f, 0.f};
      sumR[a][b] = f4{0.f, 0.f, 0.f, 0.f};
      sumI[a][b] = f4{0.f, 0.f, 0.f, 0.f};
    }
  fetch(0);
  stage(0);
  __syncthreads();
  for (int c = 0; c < npc; c++) {
    const int buf = c & 1;
    const bool more = c + 1 < npc;
    if (more) fetch(c + 1);
    h8 brh[2], brl[2], bih[2], bil[2];
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int J = 32 * wc + 16 * b + li;
      brh[b] = Bs[buf][4 * lk + 0][J];
      brl[b] = Bs[buf][4 * lk + 1][J];
      bih[b] = Bs[buf][4 * lk + 2][J];
      bil[b] = Bs[buf][4 * lk + 3][J];
    }
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const int Rw = 64 * wr + 16 * a + li;
      const h8 arh = As[buf][4 * lk + 0][Rw];
      const h8 arl = As[buf][4 * lk + 1][Rw];
      const h8 aih = As[buf][4 * lk + 2][Rw];
      const h8 ail = As[buf][4 * lk + 3][Rw];
      const h8 naih = neg(aih), nail = neg(ail);
#pragma unroll
      for (int b = 0; b < 2; b++) {
        f4 r = accR[a][b], m = accI[a][b];
        r = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh, brh[b], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh, bih[b], m, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh, brl[b], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh, bil[b], m, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_16x16x32_f16(arl, brh[b], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_f16(arl, bih[b], m, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih, bih[b], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_f16(naih, brh[b], m, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih, bil[b], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_f16(naih, brl[b], m, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_16x16x32_f16(ail, bih[b], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_f16(nail, brh[b], m, 0, 0, 0);
        accR[a][b] = r;
        accI[a][b] = m;
      }
    }
    if ((c + 1) % FOLD == 0 || !more) {
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
          sumR[a][b] += accR[a][b];
          sumI[a][b] += accI[a][b];
          accR[a][b] = f4{0.f, 0.f, 0.f, 0.f};
          accI[a][b] = f4{0.f, 0.f, 0.f, 0.f};
        }
    }
    if (more) stage(buf ^ 1);
    __syncthreads();
  }
  // unscale (exact) and store: Re = sumR, Im = -sumI
  const long rb = r0 + 64 * wr + 4 * lk;
  const int jb = c0 + 32 * wc + li;
#pragma unroll
  for (int b = 0; b < 2; b++) {
    const int J = jb + 16 * b;
    if (J >= ms) continue;
    const double fj = 1.0 / (double)cs[J];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const long r = rb + 16 * a + v;
        if (r >= nrows) continue;
        const double f = fj / (double)rs[r];
        const double xr = (double)sumR[a][b][v] * f, xi = -(double)sumI[a][b][v] * f;
        if (C64)
          ((float2*)out)[(size_t)r * ms + J] = float2{(float)xr, (float)xi};
        else
          ((double2*)out)[(size_t)r * ms + J] = double2{xr, xi};
      }
  }
}

extern "C" int pygw_minm_vhalf_split_basis_bytes(int mbsize, int matsize, size_t* bytes) {
  if (mbsize <= 0 || matsize <= 0 || !bytes) return PYGW_E_ARG;
  *bytes = vh::basis_bytes(mbsize, matsize);
  return PYGW_OK;
}

extern "C" int pygw_minm_vhalf_split_basis(const void* barcvm, int mbsize, int ld_barcvm, int matsize, void* basis,
                                           size_t bytes, void* stream_) {
  using namespace vh;
  hipStream_t stream = (hipStream_t)stream_;
  if (!barcvm || !basis || mbsize <= 0 || matsize <= 0 || ld_barcvm < mbsize) return PYGW_E_ARG;
  if (bytes < basis_bytes(mbsize, matsize)) return PYGW_E_ARG;
  float* cs = (float*)basis;
  h8* SB = (h8*)((char*)basis + al256(sizeof(float) * (size_t)matsize));
  const int npc = npc_of(mbsize), colp = pad_of(matsize);
  // column m of barcvm: mbsize values at barcvm + m ld (Fortran order)
  hipLaunchKernelGGL(pygw_vh_scale, dim3(matsize), dim3(256), 0, stream, (const double2*)barcvm, (long)ld_barcvm, 1L,
                     mbsize, cs);
  hipLaunchKernelGGL(pygw_vh_split, dim3((unsigned)(colp / 64), (unsigned)ceil_div(npc * 32, 64)), dim3(256), 0, stream,
                     (const double2*)barcvm, (long)matsize, (long)ld_barcvm, mbsize, colp, npc, (const float*)cs, SB);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

extern "C" int pygw_minm_vhalf_mixed_workspace(long n_rows, int n_in, size_t* bytes) {
  if (n_rows < 0 || n_in <= 0 || !bytes) return PYGW_E_ARG;
  *bytes = vh::rows_bytes(n_rows, n_in);
  return PYGW_OK;
}

extern "C" int pygw_minm_vhalf_mixed(const void* mmat, long n_rows, int n_in, long ld_mmat, const void* basis,
                                     int mbsize, int matsize, int complex64_out, void* minm, void* workspace,
                                     size_t ws_bytes, void* stream_) {
  using namespace vh;
  hipStream_t stream = (hipStream_t)stream_;
  if (!mmat || !basis || !minm || !workspace || n_rows < 0 || n_in <= 0 || n_in > mbsize || ld_mmat < n_in ||
      matsize <= 0)
    return PYGW_E_ARG;
  if (n_rows == 0) return PYGW_OK;
  if (n_rows > (1L << 30) || ws_bytes < rows_bytes(n_rows, n_in)) return PYGW_E_ARG;
  float* rs = (float*)workspace;
  h8* SA = (h8*)((char*)workspace + al256(sizeof(float) * (size_t)n_rows));
  const float* cs = (const float*)basis;
  const h8* SB = (const h8*)((const char*)basis + al256(sizeof(float) * (size_t)matsize));
  const int npc = npc_of(n_in), rowp = pad_of((int)n_rows), colp = pad_of(matsize);
  if ((size_t)rowp * 16 * 16 > 0x7fffffffUL) return PYGW_E_ARG;   // buffer-resource range of one chunk
  hipLaunchKernelGGL(pygw_vh_scale, dim3((unsigned)n_rows), dim3(256), 0, stream, (const double2*)mmat, ld_mmat, 1L,
                     n_in, rs);
  hipLaunchKernelGGL(pygw_vh_split, dim3((unsigned)(rowp / 64), (unsigned)ceil_div(npc * 32, 64)), dim3(256), 0, stream,
                     (const double2*)mmat, n_rows, ld_mmat, n_in, rowp, npc, (const float*)rs, SA);
  const int nct = colp / TB;
  const long items = (long)(rowp / TB) * nct, per = (items + 7) / 8;
  if (8 * per > 0x7fffffffL) return PYGW_E_ARG;
  if (complex64_out)
    hipLaunchKernelGGL(pygw_vh_mixed<true>, dim3((unsigned)(8 * per)), dim3(NT), 0, stream, (const h8*)SA, SB, n_rows,
                       matsize, rowp, colp, npc, (const float*)rs, cs, nct, items, per, minm);
  else
    hipLaunchKernelGGL(pygw_vh_mixed<false>, dim3((unsigned)(8 * per)), dim3(NT), 0, stream, (const h8*)SA, SB,
                       n_rows, matsize, rowp, colp, npc, (const float*)rs, cs, nct, items, per, minm);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}
