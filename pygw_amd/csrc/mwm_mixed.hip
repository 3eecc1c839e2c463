// Split-precision MWM of calc_selfc (pygw.py:3353-3358) on gfx950 MFMA
// (v_mfma_f32_16x16x32_f16):
//
//   mwm[w][r] = wkq * sum_J ( sum_I conj(M[r,I]) W[I,J,w] ) * M[r,J]
//
// Operands are split once into FP16 hi/lo fragment planes (the layout of the
// split-precision chi0 kernel): the rows of M scaled per row by 2^(14-e_r),
// each frequency's W scaled per column J by 2^(14-e_J,w), |x| < 2^14, so the
// low halves stay in FP16's normal range.  A real product takes three MFMA
// terms (hi hi + hi lo + lo hi);
//   Re X = m_r W_r + m_i W_i,   Im X = m_r W_i + (-m_i) W_r.
// One block forms one 128 x 128 tile of X = conj(M) W_w (128 rows, 128 J)
// over the whole I sum (PB/32 chunks) in FP32 accumulators; the epilogue
// unscales the tile exactly, takes it times M[r,J] in FP64 and writes the
// tile's per-row FP64 sum over its J into a partial-sum slot, and
// pygw_mwm_sum adds a row's slots in J-tile order (deterministic, no
// atomics).  Work items (w, row tile, J tile) are grouped so that the 32
// blocks an XCD runs at once share one frequency and R row tiles x all J
// tiles (R x T = 32 for T J tiles): each I chunk of a row tile's planes is
// read from HBM once and used by the T blocks of its J tiles, each chunk of
// the frequency's W planes once and used by the R blocks of its rows (blocks
// are dealt to the 8 XCDs round robin and each XCD starts its own in order).
#include "common.h"

#pragma clang fp contract(off)

typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

namespace mm {
constexpr int TB = 128;  // rows per block, J tile
constexpr int KC = 32;   // I per chunk = MFMA K
constexpr int NT = 512;  // threads per block
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
}  // namespace mm

// Row planes: SA[((c*4 + g)*4 + q)*rowp + r] holds, for row r of M scaled by
// rs[r], the 8 FP16 values (plane q: re_hi, re_lo, im_hi, im_lo) of
// I = 32c + 8g + j, j = 0..7.  Rows >= nrows and I >= PB are zero.
__global__ void __launch_bounds__(256) pygw_mwm_split_rows(const float2* __restrict__ M, int nrows, int pb, int rowp,
                                                           int npc, const float* __restrict__ rs,
                                                           h8* __restrict__ SA) {
  const long n = (long)npc * 4 * rowp;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int r = (int)(e % rowp);
    const long cg = e / rowp;  // c*4 + g
    const int I0 = (int)cg * 8;
    const float s = r < nrows ? rs[r] : 0.f;
    h8 rh, rl, ih, il;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      float2 v = float2{0.f, 0.f};
      if (r < nrows && I0 + j < pb) v = M[(size_t)r * pb + I0 + j];
      _Float16 h, l;
      mm::split(v.x * s, h, l);
      rh[j] = h;
      rl[j] = l;
      mm::split(v.y * s, h, l);
      ih[j] = h;
      il[j] = l;
    }
    h8* o = SA + (size_t)cg * 4 * rowp + r;
    o[0] = rh;
    o[(size_t)rowp] = rl;
    o[2 * (size_t)rowp] = ih;
    o[3 * (size_t)rowp] = il;
  }
}

// W planes for frequencies w0..w0+nw-1: SB[(((wl*npc + c)*4 + g)*4 + q)*pbp + J]
// holds W[I, J, w0+wl] * cs[w][J] for I = 32c + 8g + j (W: Fortran
// [PB, PB, n_om] complex128, so the 8 values are contiguous).
__global__ void __launch_bounds__(256) pygw_mwm_split_w(const double2* __restrict__ W, int pb, int pbp, int npc,
                                                        int w0, int nw, const float* __restrict__ cs,
                                                        h8* __restrict__ SB) {
  const long n = (long)nw * npc * 4 * pbp;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int J = (int)(e % pbp);
    const long t = e / pbp;  // (wl*npc + c)*4 + g
    const int wl = (int)(t / (4L * npc));
    const int I0 = (int)(t % (4L * npc)) * 8;
    const int w = w0 + wl;
    const double s = J < pb ? (double)cs[(size_t)w * pb + J] : 0.0;
    const double2* col = W + (size_t)w * pb * pb + (size_t)pb * (J < pb ? J : 0);
    h8 rh, rl, ih, il;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      double2 v = double2{0.0, 0.0};
      if (J < pb && I0 + j < pb) v = col[I0 + j];
      _Float16 h, l;
      mm::split((float)(v.x * s), h, l);
      rh[j] = h;
      rl[j] = l;
      mm::split((float)(v.y * s), h, l);
      ih[j] = h;
      il[j] = l;
    }
    h8* o = SB + (size_t)t * 4 * pbp + J;
    o[0] = rh;
    o[(size_t)pbp] = rl;
    o[2 * (size_t)pbp] = ih;
    o[3 * (size_t)pbp] = il;
  }
}

// rs[r] = 2^(14-e), e the exponent of max_I max(|Re|,|Im|) of row r
__global__ void __launch_bounds__(256) pygw_mwm_rowscale(const float2* __restrict__ M, int nrows, int pb,
                                                         float* __restrict__ rs) {
  const int row = blockIdx.x;
  float m = 0.f;
  for (int I = threadIdx.x; I < pb; I += blockDim.x) {
    const float2 v = M[(size_t)row * pb + I];
    m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
  }
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  __shared__ float red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    int e = 0;
    if (m > 0.f) frexpf(m, &e);
    rs[row] = ldexpf(1.0f, 14 - e);
  }
}

// cs[w][J] = 2^(14-e), e the exponent of max_I max(|Re|,|Im|) of column J of W_w
__global__ void __launch_bounds__(64) pygw_mwm_colscale(const double2* __restrict__ W, int pb, float* __restrict__ cs) {
  const int J = blockIdx.x, w = blockIdx.y;
  const double2* col = W + (size_t)w * pb * pb + (size_t)pb * J;
  double m = 0.0;
  for (int I = threadIdx.x; I < pb; I += blockDim.x) m = fmax(m, fmax(fabs(col[I].x), fabs(col[I].y)));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if (threadIdx.x == 0) {
    int e = 0;
    if (m > 0.0) frexp(m, &e);
    cs[(size_t)w * pb + J] = (float)ldexp(1.0, 14 - e);
  }
}

// MFMA 16x16x32 f16 layouts: A[i][k]: i = l%16, k = 8*(l/16) + j;
// B[k][n]: n = l%16, k = 8*(l/16) + j;  D[i][n]: n = l%16, i = 4*(l/16) + v.
// part: [nw][T][rowp] complex128, the row sums of each (frequency, J tile).
// Block b runs on XCD b % 8 as its (b / 8)-th block: group g = (b / 8) /
// (R T) of that XCD is supergroup 8 g + b % 8 = (frequency, row group of R
// row tiles), and within it item (b / 8) % (R T) = (J tile, row tile).
__global__ void __launch_bounds__(512) pygw_mwm_mixed(const h8* __restrict__ SA, const h8* __restrict__ SB,
                                                      const float2* __restrict__ M, int nrows, int pb, int rowp,
                                                      int pbp, int npc, const float* __restrict__ rs,
                                                      const float* __restrict__ cs, int w0, int nw, int nrt, int R,
                                                      double2* __restrict__ part) {
  using namespace mm;
  __shared__ h8 As[2][16][TB];  // [buf][g*4 + plane][row]
  __shared__ h8 Bs[2][16][TB];  // [buf][g*4 + plane][J]
  __shared__ double2 racc[4][TB];
  const int T = pbp / TB;
  const int ngrp = (nrt + R - 1) / R;      // row groups per frequency
  const int local = blockIdx.x >> 3, per = R * T;
  const int sg = (local / per) * 8 + (blockIdx.x & 7);
  const int within = local % per;
  const int wl = sg / ngrp, rg = sg % ngrp;
  const int jt = within / R, rt = rg * R + within % R;
  if (wl >= nw || rt >= nrt) return;       // the whole block leaves before any barrier
  const int w = w0 + wl;
  const int r0 = rt * TB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 2, wc = wave & 3;
  const int li = lane & 15, lk = lane >> 4;
  const float* csw = cs + (size_t)w * pb;

  // staging: 4 x 16 B of each slab per thread, planes contiguous in memory
  const unsigned aoffB = (unsigned)((tid >> 7) * rowp + (tid & 127)) * 16u;
  const unsigned boffB = (unsigned)((tid >> 7) * pbp + (tid & 127)) * 16u;
  const unsigned aqB = 4u * (unsigned)rowp * 16u, bqB = 4u * (unsigned)pbp * 16u;
  h8 pa[4], pbv[4];
  auto fetch = [&](int c) {
    const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(SA + (size_t)c * 16 * rowp + r0), (short)0, (int)(16u * (unsigned)rowp * 16u), RSRC_FLAGS);
    const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(SB + ((size_t)wl * npc + c) * 16 * pbp + (size_t)jt * TB), (short)0, (int)(16u * (unsigned)pbp * 16u),
        RSRC_FLAGS);
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

  f4 accR[4][2], accI[4][2];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      accR[a][b] = f4{0.f, 0.f, 0.f, 0.f};
      accI[a][b] = f4{0.f, 0.f, 0.f, 0.f};
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
    if (more) stage(buf ^ 1);
    __syncthreads();
  }
  // the X tile (rows x J) times M[r, J], summed over the tile's J per row
  double2 rsum[4][4];
  {
    const int rb = r0 + 64 * wr + 4 * lk, jb = jt * TB + 32 * wc + li;
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int v = 0; v < 4; v++) rsum[a][v] = double2{0.0, 0.0};
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int J = jb + 16 * b;
      if (J >= pb) continue;
      const double fj = 1.0 / (double)csw[J];
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int r = rb + 16 * a + v;
          if (r < nrows) {
            const double f = fj / (double)rs[r];  // exact: powers of two
            const double xr = (double)accR[a][b][v] * f, xi = (double)accI[a][b][v] * f;
            const float2 m = M[(size_t)r * pb + J];
            rsum[a][v].x += xr * (double)m.x - xi * (double)m.y;
            rsum[a][v].y += xr * (double)m.y + xi * (double)m.x;
          }
        }
    }
  }
  // reduce the row sums over the 16 lanes of a row group, then over the 4 J waves
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      double x = rsum[a][v].x, y = rsum[a][v].y;
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) {
        x += __shfl_xor(x, off, 64);
        y += __shfl_xor(y, off, 64);
      }
      if (li == 0) racc[wc][64 * wr + 16 * a + 4 * lk + v] = double2{x, y};
    }
  __syncthreads();
  for (int r = tid; r < TB; r += NT) {
    double sx = 0.0, sy = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) { sx += racc[q][r].x; sy += racc[q][r].y; }
    part[((size_t)wl * T + jt) * rowp + r0 + r] = double2{sx, sy};
  }
}

// out[k][w0 + wl][rr] = wkq * sum over the J tiles (in order) of part[wl][jt][row]
__global__ void pygw_mwm_sum(const double2* __restrict__ part, int nrows, int rowp, int T, int w0, int nw, double wkq,
                             int rpk, int nom, double2* __restrict__ out) {
  const long n = (long)nw * nrows;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int wl = (int)(e / nrows), row = (int)(e % nrows);
    const double2* p = part + (size_t)wl * T * rowp + row;
    double sx = 0.0, sy = 0.0;
    for (int jt = 0; jt < T; jt++) {
      sx += p[(size_t)jt * rowp].x;
      sy += p[(size_t)jt * rowp].y;
    }
    const int k = row / rpk, rr = row % rpk;
    out[((size_t)k * nom + w0 + wl) * rpk + rr] = double2{wkq * sx, wkq * sy};
  }
}

namespace {
size_t align256(size_t x) { return (x + 255) / 256 * 256; }
size_t rows_planes_bytes(int nrows, int pb) {
  const size_t npc = (size_t)(pb + mm::KC - 1) / mm::KC;
  const size_t rowp = (size_t)(nrows + mm::TB - 1) / mm::TB * mm::TB;
  return npc * 16 * rowp * sizeof(h8);
}
size_t w_planes_bytes_per_w(int pb) {
  const size_t npc = (size_t)(pb + mm::KC - 1) / mm::KC;
  const size_t pbp = (size_t)(pb + mm::TB - 1) / mm::TB * mm::TB;
  return npc * 16 * pbp * sizeof(h8);
}
// per frequency: its W planes and the row sums of its J tiles
size_t per_w_bytes(int nrows, int pb) {
  const size_t T = (size_t)(pb + mm::TB - 1) / mm::TB;
  const size_t rowp = (size_t)(nrows + mm::TB - 1) / mm::TB * mm::TB;
  return w_planes_bytes_per_w(pb) + align256(T * rowp * sizeof(double2));
}
size_t head_bytes(int nrows, int pb, int nom) {
  return align256(sizeof(float) * (size_t)nrows) + align256(sizeof(float) * (size_t)pb * nom);
}
constexpr size_t kWBudget = (size_t)2 << 30;  // W planes and row sums per slab of frequencies, bytes
}  // namespace

// Workspace: scales, the row planes, and the W planes and row sums of a slab
// of min(n_om, 2 GiB / per-frequency bytes) frequencies.
size_t mwm_mixed_ws_bytes(int nrows, int pb, int nom) {
  const size_t per_w = per_w_bytes(nrows, pb);
  size_t ws = kWBudget / per_w;
  if (ws < 1) ws = 1;
  if (ws > (size_t)nom) ws = (size_t)nom;
  return head_bytes(nrows, pb, nom) + rows_planes_bytes(nrows, pb) + ws * per_w;
}

int mwm_mixed_launch(const void* M, int nrows, int pb, const void* W, int nom, double wkq, int rpk, void* ws,
                     size_t ws_bytes, void* out, hipStream_t stream) {
  using namespace mm;
  const size_t head = head_bytes(nrows, pb, nom), rbytes = rows_planes_bytes(nrows, pb);
  const size_t per_w = per_w_bytes(nrows, pb), wpl = w_planes_bytes_per_w(pb);
  if (ws_bytes < head + rbytes + per_w) return PYGW_E_ARG;
  size_t wslab = (ws_bytes - head - rbytes) / per_w;
  if (wslab > (size_t)nom) wslab = (size_t)nom;
  float* rs = (float*)ws;
  float* cs = (float*)((char*)ws + align256(sizeof(float) * (size_t)nrows));
  h8* SA = (h8*)((char*)ws + head);
  h8* SB = (h8*)((char*)ws + head + rbytes);
  double2* part = (double2*)((char*)ws + head + rbytes + wslab * wpl);  // wpl: a multiple of 256 B
  const int npc = ceil_div(pb, KC);
  const int rowp = ceil_div(nrows, TB) * TB;
  const int pbp = ceil_div(pb, TB) * TB;
  const int nrt = rowp / TB, T = pbp / TB;
  // row tiles per XCD group: R x T blocks of one frequency run together
  int R = 32 / T;
  if (R < 1) R = 1;
  if (R > nrt) R = nrt;
  const int ngrp = ceil_div(nrt, R);
  hipLaunchKernelGGL(pygw_mwm_rowscale, dim3(nrows), dim3(256), 0, stream, (const float2*)M, nrows, pb, rs);
  hipLaunchKernelGGL(pygw_mwm_colscale, dim3(pb, nom), dim3(64), 0, stream, (const double2*)W, pb, cs);
  {
    const long n = (long)npc * 4 * rowp;
    int blocks = (int)((n + 255) / 256);
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(pygw_mwm_split_rows, dim3(blocks), dim3(256), 0, stream, (const float2*)M, nrows, pb, rowp, npc,
                       (const float*)rs, SA);
  }
  for (int w0 = 0; w0 < nom; w0 += (int)wslab) {
    const int nw = (nom - w0) < (int)wslab ? (nom - w0) : (int)wslab;
    const long n = (long)nw * npc * 4 * pbp;
    int blocks = (int)((n + 255) / 256);
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(pygw_mwm_split_w, dim3(blocks), dim3(256), 0, stream, (const double2*)W, pb, pbp, npc, w0, nw,
                       (const float*)cs, SB);
    const long nsg = (long)nw * ngrp;             // supergroups (frequency, row group)
    const long grid = ((nsg + 7) / 8) * 8 * R * T;
    if (grid > 0x7fffffffL) return PYGW_E_ARG;
    hipLaunchKernelGGL(pygw_mwm_mixed, dim3((unsigned)grid), dim3(NT), 0, stream, (const h8*)SA, (const h8*)SB,
                       (const float2*)M, nrows, pb, rowp, pbp, npc, (const float*)rs, (const float*)cs, w0, nw, nrt, R,
                       part);
    const long ne = (long)nw * nrows;
    int sblocks = (int)((ne + 255) / 256);
    if (sblocks > 65536) sblocks = 65536;
    hipLaunchKernelGGL(pygw_mwm_sum, dim3(sblocks), dim3(256), 0, stream, (const double2*)part, nrows, rowp, T, w0, nw,
                       wkq, rpk, nom, (double2*)out);
  }
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}
