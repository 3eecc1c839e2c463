// Split-precision chi0 contraction on gfx950 MFMA (v_mfma_f32_16x16x32_f16).
//
//   eps[I,J,w] += sum_k sum_p coef*kcw[k,w,p] * M[k,p,I] * conj(M[k,p,J])
//
// Operands.  M (complex64) is scaled per column by an exact power of two s_I
// (|Re|,|Im| < 2^14, high in FP16's normal range so that the low halves of
// small entries stay normal) and the weights per frequency by a power of two
// t_w (|coef*kcw*t_w| < 1).  A real operand x is split into FP16
// x_hi = fp16(x), x_lo = fp16(x - x_hi) (22 significant bits) and a real
// product uses three MFMA terms a_hi b_hi + a_hi b_lo + a_lo b_hi.
//   I side:  a = s_I M[p,I], split once per call by pygw_chi0_presplit into
//            four FP16 planes in MFMA fragment order (frequency independent).
//   J side:  b = (s_J M[p,J]) (w_p t_w), weighted and split per frequency
//            while it is staged into LDS (one block = one frequency).
//   Re += a_r b_r + a_i b_i,   Im += a_i b_r + (-a_r) b_i   (12 MFMAs per
//   32x32 tile and 16 pairs on v_mfma_f32_32x32x16_f16; -a_r by a sign flip
//   of the FP16 planes).  The 32x32 shape keeps the MFMA pipe full with the
//   four independent accumulator chains a wave holds (2 tiles x Re/Im);
//   16x16x32 needs ~16 chains for the same rate (measured).
//
// Accumulation.  The MFMA accumulators (FP32; the accumulate rounds to
// nearest, measured on gfx950) carry G chunks of 32 pairs; then the tile is
// unscaled by 1/(s_I s_J t_w) (exact) and added into the FP64 eps in memory,
// so the long K = N_k*N_p reduction is carried in FP64.
//
// Tiling.  512 threads = 8 waves per block, block tile 128 (I) x 128 (J),
// wave tile 64 x 32 (2 x 4 waves; two 32x32 MFMA tiles), two waves per SIMD.  LDS holds the staged
// I and J slabs of the current and the next chunk (2 x 2 x 32 KB); the next
// chunk's global loads are issued before the current chunk's MFMAs.  Only
// tile pairs I_t <= J_t are accumulated, stored conjugated at the transposed
// place (coalesced); pygw_chi0_mirror fills the rest.  The
// block -> (tile pair, frequency) map keeps all frequencies of one tile pair
// on one XCD (blocks are dealt to the 8 XCDs round robin), so the slabs they
// share are read from HBM once per XCD and then served from its L2.
#include "common.h"

#include <algorithm>
#include <vector>

// No FMA contraction in this file, and every split goes through cm::split,
// which pins x to a rounded f32 register first: the hi/lo split of x = a*b
// must use the same f32 product in x_hi = fp16(x) and x_lo = fp16(x - x_hi).
// Otherwise the compiler may form x_hi = fp16(a*b) in one rounding
// (v_fma_mixlo_f16) and x_lo from the f32-rounded product; for products on an
// f16 rounding tie the pair then misses x by a full ulp of x_lo (measured on
// gfx950: 1e-5 relative errors in whole chi0 columns).
#pragma clang fp contract(off)

typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

namespace cm {
constexpr int TB = 128;  // block tile edge (I and J)
constexpr int KC = 32;   // pairs per chunk = MFMA K
constexpr int NT = 512;  // threads per block

__device__ __forceinline__ void tile_pair(int t, int T, int& It, int& Jt) {
  int i = 0;
  while (t >= T - i) { t -= T - i; i++; }
  It = i;
  Jt = i + t;
}

__device__ __forceinline__ h8 neg(h8 x) {
  u4 u = __builtin_bit_cast(u4, x);
  u ^= u4{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
  return __builtin_bit_cast(h8, u);
}

// x = hi + lo to 22 bits; the empty asm makes x an opaque f32 value so both
// halves are taken from the same rounding of it
__device__ __forceinline__ void split(float x, _Float16& hi, _Float16& lo) {
  asm volatile("" : "+v"(x));
  hi = (_Float16)x;
  lo = (_Float16)(x - (float)hi);
}
}  // namespace cm

// Fragment planes.  For k-point k (of the slab), chunk c, group g = 0..3 and
// plane q (0 re_hi, 1 re_lo, 2 im_hi, 3 im_lo), column I holds the 8 FP16
// values of pairs p = 32c + 8g + j, j = 0..7 (16 bytes):
//   S[(((k*npc + c)*4 + g)*4 + q)*pbp + I]     (h8 units; pbp = PB padded to 128)
// Columns I >= PB and pairs p >= N_p are zero.
__global__ void __launch_bounds__(256) pygw_chi0_presplit(const float2* __restrict__ M, long strideK, int np, int pb,
                                                          int pbp, int npc, int nkc, const float* __restrict__ cs,
                                                          h8* __restrict__ S) {
  const long n = (long)nkc * npc * 4 * pbp;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int I = (int)(e % pbp);
    long r = e / pbp;
    const int g = (int)(r & 3);
    r >>= 2;
    const int c = (int)(r % npc);
    const int k = (int)(r / npc);
    const float s = I < pb ? cs[I] : 0.f;
    h8 rh, rl, ih, il;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int p = 32 * c + 8 * g + j;
      float2 v = float2{0.f, 0.f};
      if (p < np && I < pb) v = M[(size_t)k * strideK + (size_t)p * pb + I];
      _Float16 h, l;
      cm::split(v.x * s, h, l);
      rh[j] = h;
      rl[j] = l;
      cm::split(v.y * s, h, l);
      ih[j] = h;
      il[j] = l;
    }
    h8* o = S + (size_t)(e / pbp) * 4 * pbp + I;  // ((k*npc + c)*4 + g)*4*pbp + I
    o[0] = rh;
    o[(size_t)pbp] = rl;
    o[2 * (size_t)pbp] = ih;
    o[3 * (size_t)pbp] = il;
  }
}

// MFMA 32x32x16 f16 operand layouts (lane l, element j = 0..7 of the h8;
// probed on gfx950):
//   A[i][k]: i = l%32, k = 8*(l/32) + j;  B[k][n]: n = l%32, k = 8*(l/32) + j;
//   D[i][n] (16 values v): n = l%32, i = 8*(v/4) + 4*(l/32) + v%4.
// S: fragment planes of this slab; M: the same slab as complex64 (J side);
// wf: the slab's scaled weights coef*kcw*t_w in FP32, [k][w][npc*32] (zero
// past N_p), from pygw_chi0_wprep.
// One (tile pair, frequency) work item: eps[.., om] += the 128 x 128 tile
// (It, Jt) of the contraction over all chunks of S / M / wf.  As and Bs are
// the block's LDS buffers; every thread of the 512-thread block calls this
// with the same arguments, and it ends on a block barrier.
__device__ __forceinline__ void chi0_tile(const h8* __restrict__ S, const float2* __restrict__ M, long strideK,
                                          int np, int pb, int pbp, int npc, int nkc, const float* __restrict__ wf,
                                          int nom, const float* __restrict__ cs, const double* __restrict__ ws,
                                          int G, int om, int It, int Jt, double2* __restrict__ eps,
                                          h8 (*As)[16][cm::TB], h8 (*Bs)[16][cm::TB]) {
  using namespace cm;
  const int I0 = It * TB, J0 = Jt * TB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 2, wc = wave & 3;
  const double tw = ws[om], itw = 1.0 / tw;  // exact: a power of two
  const float* ics = cs + pb;                  // 1/cs, exact

  // staging roles: A: 4 x 16 B per thread (contiguous copy of 16 planes x 128
  // columns); B: column Jb = tid % 128, pair group gb = tid / 128 (8 pairs)
  const int Jb = tid & 127, gb = tid >> 7;
  const bool jin = (J0 + Jb) < pb;
  const float sJ = jin ? cs[J0 + Jb] : 0.f;
  // one chunk's staged inputs, held in registers between its load and its
  // LDS store (two sets: the next chunk is stored while the one after it
  // loads)
  struct Pre {
    h8 a[4];
    float2 m[8];
    float w[8];
  };
  const long nch = (long)nkc * npc;
  // Loads go through buffer resources built from chunk-uniform bases (scalar
  // registers) with per-lane byte offsets that stay fixed across chunks, and
  // out-of-range offsets read as zero: J-side pairs p >= N_p fall past the
  // end of the k-point's M block.
  constexpr int RSRC_FLAGS = 0x00020000;
  const unsigned aoffB = (unsigned)((tid >> 7) * pbp + (tid & 127)) * 16u;  // + q * 4*pbp*16
  const unsigned aqB = 4u * (unsigned)pbp * 16u;
  const unsigned pb8 = (unsigned)pb * 8u;
  // columns past PB read column PB-1 (finite) and are zeroed by sJ = 0
  const unsigned moff0 = (unsigned)(8 * gb) * pb8 + (unsigned)(jin ? J0 + Jb : pb - 1) * 8u;
  const int npad = npc * KC;
  auto fetch = [&](Pre& P, int k, int c) {
    const h8* Sk = S + ((size_t)k * npc + c) * 16 * (size_t)pbp + I0;
    const __amdgpu_buffer_rsrc_t ra =
        __builtin_amdgcn_make_buffer_rsrc((void*)Sk, (short)0, (int)(16u * (unsigned)pbp * 16u), RSRC_FLAGS);
#pragma unroll
    for (int q = 0; q < 4; q++)
      P.a[q] = __builtin_bit_cast(h8, __builtin_amdgcn_raw_buffer_load_b128(ra, aoffB + q * aqB, 0, 0));
    const __amdgpu_buffer_rsrc_t rm = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(M + (size_t)k * strideK), (short)0, (int)((unsigned)np * pb8), RSRC_FLAGS);
    const unsigned mo = moff0 + (unsigned)(32 * c) * pb8;
#pragma unroll
    for (int j = 0; j < 8; j++)
      P.m[j] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rm, mo + j * pb8, 0, 0));
    const __amdgpu_buffer_rsrc_t rw = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(wf + ((size_t)k * nom + om) * (size_t)npad), (short)0, npad * 4, RSRC_FLAGS);
    const unsigned wo = (unsigned)(32 * c + 8 * gb) * 4u;
    const u4 w0 = __builtin_amdgcn_raw_buffer_load_b128(rw, wo, 0, 0);
    const u4 w1 = __builtin_amdgcn_raw_buffer_load_b128(rw, wo + 16, 0, 0);
    P.w[0] = __uint_as_float(w0.x); P.w[1] = __uint_as_float(w0.y); P.w[2] = __uint_as_float(w0.z); P.w[3] = __uint_as_float(w0.w);
    P.w[4] = __uint_as_float(w1.x); P.w[5] = __uint_as_float(w1.y); P.w[6] = __uint_as_float(w1.z); P.w[7] = __uint_as_float(w1.w);
  };
  auto stage = [&](const Pre& P, int buf) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int idx = tid + NT * q;
      As[buf][idx >> 7][idx & 127] = P.a[q];
    }
    h8 rh, rl, ih, il;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      _Float16 h, l;
      split((P.m[j].x * sJ) * P.w[j], h, l);
      rh[j] = h;
      rl[j] = l;
      split((P.m[j].y * sJ) * P.w[j], h, l);
      ih[j] = h;
      il[j] = l;
    }
    Bs[buf][4 * gb + 0][Jb] = rh;
    Bs[buf][4 * gb + 1][Jb] = rl;
    Bs[buf][4 * gb + 2][Jb] = ih;
    Bs[buf][4 * gb + 3][Jb] = il;
  };

  // accumulators: two 32x32 tiles (rows 64*wr + 32*a) of the wave's 64 x 32
  f16v accR[2], accI[2];
  auto zero = [&]() {
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int v = 0; v < 16; v++) {
        accR[a][v] = 0.f;
        accI[a][v] = 0.f;
      }
  };
  // The tile is added, conjugated, at its transposed place eps[J, I]: there
  // the 32 lanes of a half-wave (consecutive J, one I) touch 512 contiguous
  // bytes of the column-major eps.  pygw_chi0_mirror then fills the tiles
  // above the diagonal from the ones below.
  auto flush = [&]() {
    // opaque copies keep the compiler from hoisting the 32 element addresses
    // (and their scales) out of the chunk loop into live registers
    int ib = I0 + 64 * wr + 4 * (lane >> 5), J = J0 + 32 * wc + (lane & 31);
    asm volatile("" : "+v"(ib), "+v"(J));
    if (J >= pb) return;
    double2* E = eps + (size_t)pb * pb * om + J;
    const double fj = (double)ics[J] * itw;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const int I = ib + 32 * a + 8 * (v >> 2) + (v & 3);
        if (I < pb) {
          const double f = fj * (double)ics[I];  // exact: powers of two
          double2* e = E + (size_t)pb * I;
          const double2 o = *e;
          *e = double2{o.x + (double)accR[a][v] * f, o.y - (double)accI[a][v] * f};
        }
      }
  };

  // chunk `buf`'s products (LDS buffer buf), with the next chunk's inputs
  // from S stored into the other buffer in the same basic block so that the
  // compiler can interleave that VALU/LDS work with the MFMAs
  auto body = [&](int buf, const Pre& S) {
    // J-side fragments of the wave's 32 columns for the two K steps of 16
    const int Jw = 32 * wc + (lane & 31), hk = lane >> 5;
    h8 brh[2], brl[2], bih[2], bil[2];
#pragma unroll
    for (int st = 0; st < 2; st++) {
      const int g = 2 * st + hk;
      brh[st] = Bs[buf][4 * g + 0][Jw];
      brl[st] = Bs[buf][4 * g + 1][Jw];
      bih[st] = Bs[buf][4 * g + 2][Jw];
      bil[st] = Bs[buf][4 * g + 3][Jw];
    }
    stage(S, buf ^ 1);
#pragma unroll
    for (int st = 0; st < 2; st++) {
      const int g = 2 * st + hk;
#pragma unroll
      for (int a = 0; a < 2; a++) {
        const int I = 64 * wr + 32 * a + (lane & 31);
        const h8 arh = As[buf][4 * g + 0][I];
        const h8 arl = As[buf][4 * g + 1][I];
        const h8 aih = As[buf][4 * g + 2][I];
        const h8 ail = As[buf][4 * g + 3][I];
        const h8 narh = neg(arh), narl = neg(arl);  // Im = a_i b_r + (-a_r) b_i
        f16v r = accR[a], m = accI[a];
        r = __builtin_amdgcn_mfma_f32_32x32x16_f16(arh, brh[st], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_32x32x16_f16(aih, brh[st], m, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_32x32x16_f16(arh, brl[st], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_32x32x16_f16(aih, brl[st], m, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_32x32x16_f16(arl, brh[st], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_32x32x16_f16(ail, brh[st], m, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_32x32x16_f16(aih, bih[st], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_32x32x16_f16(narh, bih[st], m, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_32x32x16_f16(aih, bil[st], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_32x32x16_f16(narh, bil[st], m, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_32x32x16_f16(ail, bih[st], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_32x32x16_f16(narl, bih[st], m, 0, 0, 0);
        accR[a] = r;
        accI[a] = m;
      }
    }
  };

  // Pipeline: chunk ch computes from LDS buffer ch&1 while chunk ch+1 (in
  // registers since the previous step) is stored into the other buffer and
  // chunk ch+2 loads.  Past the end the last chunk is loaded and stored
  // again (harmless: that buffer is not read any more).
  const long last = nch - 1;
  int kf = 0, cf = 0;  // next chunk to fetch
  long nf = 0;         // its index
  auto fetch_next = [&](Pre& P) {
    fetch(P, kf, cf);
    if (nf < last) {
      nf++;
      if (++cf == npc) { cf = 0; kf++; }
    }
  };
  zero();
  Pre P0, P1;
  fetch_next(P0);
  stage(P0, 0);
  fetch_next(P1);
  __syncthreads();
  int g = 0;
  for (long ch = 0; ch < nch; ch += 2) {
    fetch_next(P0);  // chunk ch+2
    body(0, P1);     // computes ch, stores ch+1
    if (++g == G || ch == last) {
      flush();
      zero();
      g = 0;
    }
    __syncthreads();
    if (ch == last) break;
    fetch_next(P1);  // chunk ch+3
    body(1, P0);     // computes ch+1, stores ch+2
    if (++g == G || ch + 1 == last) {
      flush();
      zero();
      g = 0;
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(512) pygw_chi0_mixed(const h8* __restrict__ S, const float2* __restrict__ M,
                                                       long strideK, int np, int pb, int pbp, int npc, int nkc,
                                                       const float* __restrict__ wf, int nom,
                                                       const float* __restrict__ cs,
                                                       const double* __restrict__ ws, int G, int T, int nwork,
                                                       double2* __restrict__ eps) {
  using namespace cm;
  __shared__ h8 As[2][16][TB];  // [buf][g*4 + plane][I]
  __shared__ h8 Bs[2][16][TB];  // [buf][g*4 + plane][J], weighted
  // XCD-aware work map: blocks b, b+8, b+16, ... run on one XCD and take
  // consecutive work items (frequency fastest).
  const int per = gridDim.x >> 3;
  const int work = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (work >= nwork) return;
  int It, Jt;
  tile_pair(work / nom, T, It, Jt);
  chi0_tile(S, M, strideK, np, pb, pbp, npc, nkc, wf, nom, cs, ws, G, work % nom, It, Jt, eps, As, Bs);
}

// Tiles above the diagonal from the ones below: eps[I,J] = conj(eps[J,I])
// for I in tile It < Jt (the contraction writes the tiles It >= Jt).  One
// thread per element.
__global__ void pygw_chi0_mirror(double2* __restrict__ eps, int pb, int nom) {
  const long n = (long)pb * pb;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int I = (int)(e % pb), J = (int)(e / pb);  // element (I, J), column-major
  if ((I / cm::TB) >= (J / cm::TB)) return;         // computed tile
  double2* E = eps + (size_t)blockIdx.y * n;
  const double2 v = E[(size_t)J + (size_t)pb * I];
  E[(size_t)I + (size_t)pb * J] = double2{v.x, -v.y};
}

// Per-column maximum of max(|Re|,|Im|) over all (k, p): one block per
// 64-column strip and k-range, a block reduction then one atomic per column.
__global__ void __launch_bounds__(256) pygw_colmax(const float2* __restrict__ M, long strideK, int np, int pb, int nk,
                                                   unsigned int* __restrict__ cmax) {
  __shared__ float red[4][64];
  const int c = threadIdx.x & 63, r = threadIdx.x >> 6;
  const int I = blockIdx.x * 64 + c;
  float m = 0.f;
  // rows (k, p) flattened, dealt 4 per block step over gridDim.y blocks
  const int step = 4 * gridDim.y;
  int k = 0, p = r + 4 * blockIdx.y;
  while (k < nk && p >= np) { p -= np; k++; }
  if (I < pb)
    while (k < nk) {
      const float2 v = M[(size_t)k * strideK + (size_t)p * pb + I];
      m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
      p += step;
      while (k < nk && p >= np) { p -= np; k++; }
    }
  red[r][c] = m;
  __syncthreads();
  if (r == 0 && I < pb) {
    m = fmaxf(fmaxf(red[0][c], red[1][c]), fmaxf(red[2][c], red[3][c]));
    atomicMax(&cmax[I], __float_as_uint(m));  // non-negative floats order as uints
  }
}

// blocks along the rows of pygw_colmax: about 64 rows per block, at most 1024
static int colmax_rows_grid(long rows) {
  const long g = (rows + 63) / 64;
  return (int)(g < 1 ? 1 : g > 1024 ? 1024 : g);
}

// cs[I] = 2^(14 - e), e the exponent of the column maximum (|x cs| < 2^14);
// cs[pb + I] = 1 / cs[I]
__global__ void pygw_colscale(const unsigned int* __restrict__ cmax, int pb, float* __restrict__ cs) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= pb) return;
  const float m = __uint_as_float(cmax[I]);
  int e = 0;
  if (m > 0.f) frexpf(m, &e);  // m = f * 2^e, f in [0.5, 1)
  cs[I] = ldexpf(1.0f, 14 - e);
  cs[pb + I] = ldexpf(1.0f, e - 14);  // its exact inverse
}

// per-frequency maximum of |coef*kcw| over (k, p): block reduction, one
// atomic per block
__global__ void __launch_bounds__(256) pygw_wmax(const double* __restrict__ kcw, int np, int nom, int nk, double coef,
                                                 unsigned long long* __restrict__ wmax) {
  __shared__ double red[4];
  const int w = blockIdx.y;
  double m = 0.0;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)nk * np; e += (long)gridDim.x * blockDim.x) {
    const int k = (int)(e / np), p = (int)(e % np);
    m = fmax(m, fabs(coef * kcw[((size_t)k * nom + w) * np + p]));
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    atomicMax(&wmax[w], (unsigned long long)__double_as_longlong(m));
  }
}

// ws[w] = 2^-e, e the exponent of the frequency maximum (|coef kcw ws| < 1)
__global__ void pygw_wscale(const unsigned long long* __restrict__ wmax, int nom, double* __restrict__ ws) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nom) return;
  const double m = __longlong_as_double((long long)wmax[w]);
  int e = 0;
  if (m > 0.0) frexp(m, &e);
  ws[w] = ldexp(1.0, -e);
}

__global__ void pygw_zero_u32(unsigned int* __restrict__ x, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) x[e] = 0u;
}

__global__ void pygw_zero_c128(double2* __restrict__ x, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    x[e] = double2{0.0, 0.0};
}

// wf[k][w][p] = (float)(coef * kcw[k0+k][w][p] * ws[w]) for p < N_p, 0 up to
// npc*32: the J-side weights of one slab, scaled and rounded once.
__global__ void pygw_chi0_wprep(const double* __restrict__ kcw, int np, int nom, int k0, int nkc, int npad, double coef,
                                const double* __restrict__ ws, float* __restrict__ wf) {
  const long n = (long)nkc * nom * npad;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int p = (int)(e % npad);
    const long kw = e / npad;  // k*nom + w
    const int w = (int)(kw % nom);
    wf[e] = p < np ? (float)(coef * kcw[((size_t)k0 * nom + kw) * np + p] * ws[w]) : 0.f;
  }
}

namespace {
constexpr size_t kHead = 1 << 16;                // scales region granule, bytes
constexpr size_t kSlabBudget = (size_t)1 << 30;  // default slab budget, bytes

size_t head_bytes(int pb, int nom) {
  const size_t h = (size_t)pb * 12 + (size_t)nom * 16 + 512;
  return (h + kHead - 1) / kHead * kHead;
}
// fragment planes + FP32 weights of one k-point (both multiples of 256 B)
size_t slab_bytes_per_k(int np, int pb, int nom) {
  const size_t npc = (size_t)(np + cm::KC - 1) / cm::KC;
  const size_t pbp = (size_t)(pb + cm::TB - 1) / cm::TB * cm::TB;
  const size_t wbytes = ((size_t)nom * npc * cm::KC * sizeof(float) + 255) / 256 * 256;
  return npc * 16 * pbp * sizeof(h8) + wbytes;
}
}  // namespace

// Workspace for a call over n_k k-points: scales + fragment planes and
// weights of up to min(n_k, 1 GiB / per-k bytes) k-points (the call walks k
// in slabs).
size_t chi0_mixed_ws_bytes(int pb, int np, int nom, int nk) {
  const size_t per_k = slab_bytes_per_k(np, pb, nom);
  size_t ks = kSlabBudget / per_k;
  if (ks < 1) ks = 1;
  if (ks > (size_t)nk) ks = (size_t)nk;
  return head_bytes(pb, nom) + ks * per_k;
}

// M: complex64 [n_kc][np][pb] holding k = k0..k0+n_kc-1; kcw covers all k.
// eps accumulates (beta = 1) or is overwritten (beta = 0).
int chi0_mixed_launch(const void* M, long strideK, int np, int pb, const double* kcw, int nom, int k0, int nkc,
                      double coef, int G, int beta, void* ws, size_t ws_bytes, void* eps, hipStream_t stream) {
  using namespace cm;
  const size_t head = head_bytes(pb, nom);
  const size_t per_k = slab_bytes_per_k(np, pb, nom);
  if (ws_bytes < head + per_k || G < 1) return PYGW_E_ARG;
  if ((size_t)(np + KC) * pb * 8 >= ((size_t)1 << 31)) return PYGW_E_ARG;  // 32-bit buffer offsets per k
  size_t ks = (ws_bytes - head) / per_k;
  if (ks > (size_t)nkc) ks = (size_t)nkc;
  float* cs = (float*)ws;  // [pb] scales, then [pb] their inverses
  unsigned int* cmax = (unsigned int*)(cs + 2 * pb);
  double* wsc = (double*)(((uintptr_t)(cmax + pb) + 15) & ~(uintptr_t)15);
  unsigned long long* wmax = (unsigned long long*)(wsc + nom);
  const int T = ceil_div(pb, TB);
  const int pbp = T * TB;
  const int npc = ceil_div(np, KC);
  h8* S = (h8*)((char*)ws + head);
  float* wf = (float*)((char*)ws + head + ks * (size_t)npc * 16 * pbp * sizeof(h8));
  const float2* Mf = (const float2*)M;

  hipLaunchKernelGGL(pygw_zero_u32, dim3(1), dim3(256), 0, stream, cmax, (long)pb);
  hipLaunchKernelGGL(pygw_zero_u32, dim3(1), dim3(256), 0, stream, (unsigned int*)wmax, 2L * nom);
  hipLaunchKernelGGL(pygw_colmax, dim3(ceil_div(pb, 64), colmax_rows_grid((long)nkc * np)), dim3(256), 0, stream, Mf,
                     strideK,
                     np, pb, nkc, cmax);
  hipLaunchKernelGGL(pygw_colscale, dim3(ceil_div(pb, 256)), dim3(256), 0, stream, cmax, pb, cs);
  hipLaunchKernelGGL(pygw_wmax, dim3(64, nom), dim3(256), 0, stream, kcw + (size_t)k0 * nom * np, np, nom, nkc, coef,
                     wmax);
  hipLaunchKernelGGL(pygw_wscale, dim3(ceil_div(nom, 256)), dim3(256), 0, stream, wmax, nom, wsc);
  const long ntot = (long)pb * pb * nom;
  if (!beta) hipLaunchKernelGGL(pygw_zero_c128, dim3(2048), dim3(256), 0, stream, (double2*)eps, ntot);

  const int nwork = T * (T + 1) / 2 * nom;
  const int grid = ceil_div(nwork, 8) * 8;
  for (long kk = 0; kk < nkc; kk += (long)ks) {
    const int n = (int)((nkc - kk) < (long)ks ? (nkc - kk) : (long)ks);
    const float2* Ms = Mf + (size_t)kk * strideK;
    const long nel = (long)n * npc * 4 * pbp;
    int blocks = (int)((nel + 255) / 256);
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(pygw_chi0_presplit, dim3(blocks), dim3(256), 0, stream, Ms, strideK, np, pb, pbp, npc, n, cs, S);
    const long nw_el = (long)n * nom * npc * KC;
    int wblocks = (int)((nw_el + 255) / 256);
    if (wblocks > 65536) wblocks = 65536;
    hipLaunchKernelGGL(pygw_chi0_wprep, dim3(wblocks), dim3(256), 0, stream, kcw, np, nom, k0 + (int)kk, n, npc * KC,
                       coef, (const double*)wsc, wf);
    hipLaunchKernelGGL(pygw_chi0_mixed, dim3(grid), dim3(NT), 0, stream, (const h8*)S, Ms, strideK, np, pb, pbp, npc,
                       n, (const float*)wf, nom, cs, wsc, G, T, nwork, (double2*)eps);
  }
  hipLaunchKernelGGL(pygw_chi0_mirror, dim3(ceil_div((long)pb * pb, 256), nom), dim3(256), 0, stream,
                     (double2*)eps, pb, nom);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

// ---------------------------------------------------------------------------
// Batched contraction over many q-points in one persistent launch.  Each q
// has its own PB and pair count (all pairs of all k concatenated: only the
// (k, p) list matters to the sum), its own operands, scales and eps.  Work
// item = (q, tile pair, frequency).  The (q, tile pair) groups, in order of
// decreasing cost (the q's chunk count), are dealt to the 8 XCDs round robin;
// within an XCD the resident blocks take the XCD's items from an atomic
// counter, frequency fastest, so blocks that run together on an XCD share a
// tile pair's operand slabs in its L2.
// ---------------------------------------------------------------------------
struct Chi0Q {
  const h8* S;        // fragment planes [npc][16][pbp]
  const float2* M;    // complex64 [K][PB]
  const float* wf;    // FP32 scaled weights [nom][npc*32]
  const float* cs;    // [2*PB] column scales and inverses
  const double* ws;   // [nom] weight scales
  double2* eps;       // Fortran [PB, PB, nom]
  int pb, pbp, np, npc;
};

__global__ void __launch_bounds__(512) pygw_chi0_mixed_batch(const Chi0Q* __restrict__ qs,
                                                             const int4* __restrict__ groups,
                                                             const int* __restrict__ glen, int nom, int G,
                                                             unsigned int* __restrict__ counters) {
  using namespace cm;
  __shared__ h8 As[2][16][TB];
  __shared__ h8 Bs[2][16][TB];
  __shared__ int item;
  const int x = blockIdx.x & 7;
  const int n = glen[x];
  for (;;) {
    if (threadIdx.x == 0) item = (int)atomicAdd(&counters[x], 1u);
    __syncthreads();
    const int i = item;
    __syncthreads();  // all threads hold i before thread 0 takes the next one
    if (i >= n) break;
    const int4 gr = groups[x + 8 * (i / nom)];  // (q, It, Jt)
    const Chi0Q d = qs[gr.x];
    chi0_tile(d.S, d.M, 0, d.np, d.pb, d.pbp, d.npc, 1, d.wf, nom, d.cs, d.ws, G, i % nom, gr.y, gr.z, d.eps,
              As, Bs);
  }
}

__global__ void pygw_zero_u32_small(unsigned int* __restrict__ x, int n) {
  if ((int)threadIdx.x < n) x[threadIdx.x] = 0u;
}

// Host-side plan of a batch (opaque to callers): per-q sizes and pointers,
// device tables in the caller's workspace.
struct Chi0BatchPlan {
  int nq, nom;
  double coef;
  std::vector<Chi0Q> q;
  std::vector<const double*> kcw;
  std::vector<size_t> off_head;  // per q, into ws
  Chi0Q* d_qs;
  int4* d_groups;
  int* d_glen;
  unsigned int* d_counters;
  int ngroups;
  char* ws;
};

namespace {
size_t batch_q_bytes(int pb, int np, int nom) {
  const size_t npc = (size_t)(np + cm::KC - 1) / cm::KC;
  const size_t pbp = (size_t)(pb + cm::TB - 1) / cm::TB * cm::TB;
  return head_bytes(pb, nom) + npc * 16 * pbp * sizeof(h8) +
         ((size_t)nom * npc * cm::KC * sizeof(float) + 255) / 256 * 256;
}
size_t batch_tables_bytes(int nq, const int* pb) {
  size_t ng = 0;
  for (int q = 0; q < nq; q++) {
    const size_t T = (size_t)(pb[q] + cm::TB - 1) / cm::TB;
    ng += T * (T + 1) / 2;
  }
  const size_t a = ((size_t)nq * sizeof(Chi0Q) + 255) / 256 * 256;
  const size_t b = ((ng + 8) * sizeof(int4) + 255) / 256 * 256;
  return a + b + 512;
}
}  // namespace

size_t chi0_batch_ws_bytes(int nq, const int* pb, const int* np, int nom) {
  size_t t = batch_tables_bytes(nq, pb);
  for (int q = 0; q < nq; q++) t += batch_q_bytes(pb[q], np[q], nom);
  return t;
}

// Builds the plan and uploads its tables (synchronously on `stream`).
int chi0_batch_plan(int nq, const void* const* M, const int* pb, const int* np, const double* const* kcw, int nom,
                    double coef, void* const* eps, void* ws, size_t ws_bytes, hipStream_t stream,
                    Chi0BatchPlan** out) {
  using namespace cm;
  if (ws_bytes < chi0_batch_ws_bytes(nq, pb, np, nom)) return PYGW_E_ARG;
  Chi0BatchPlan* P = new Chi0BatchPlan();
  P->nq = nq;
  P->nom = nom;
  P->coef = coef;
  P->ws = (char*)ws;
  size_t off = batch_tables_bytes(nq, pb);
  std::vector<int4> groups;
  for (int q = 0; q < nq; q++) {
    Chi0Q d;
    d.pb = pb[q];
    d.np = np[q];
    d.npc = ceil_div(np[q], KC);
    const int T = ceil_div(pb[q], TB);
    d.pbp = T * TB;
    char* base = (char*)ws + off;
    const size_t head = head_bytes(pb[q], nom);
    d.cs = (const float*)base;
    d.ws = (const double*)(((uintptr_t)((const unsigned int*)(d.cs + 2 * pb[q]) + pb[q]) + 15) & ~(uintptr_t)15);
    d.S = (const h8*)(base + head);
    d.wf = (const float*)(base + head + (size_t)d.npc * 16 * d.pbp * sizeof(h8));
    d.M = (const float2*)M[q];
    d.eps = (double2*)eps[q];
    P->q.push_back(d);
    P->kcw.push_back(kcw[q]);
    P->off_head.push_back(off);
    off += batch_q_bytes(pb[q], np[q], nom);
    for (int t = 0; t < T * (T + 1) / 2; t++) {
      int It, Jt;
      int tt = t, i = 0;
      while (tt >= T - i) { tt -= T - i; i++; }
      It = i;
      Jt = i + tt;
      groups.push_back(int4{q, It, Jt, 0});
    }
  }
  // longest first: a group's items cost ~ its q's chunk count, so dealing
  // them in decreasing cost leaves the short ones for the tail of each XCD
  std::stable_sort(groups.begin(), groups.end(),
                   [&](const int4& a, const int4& b) { return P->q[a.x].npc > P->q[b.x].npc; });
  P->ngroups = (int)groups.size();
  const size_t a = ((size_t)nq * sizeof(Chi0Q) + 255) / 256 * 256;
  const size_t b = (((size_t)P->ngroups + 8) * sizeof(int4) + 255) / 256 * 256;
  P->d_qs = (Chi0Q*)ws;
  P->d_groups = (int4*)((char*)ws + a);
  P->d_glen = (int*)((char*)ws + a + b);
  P->d_counters = (unsigned int*)((char*)ws + a + b + 256);
  int glen[8];
  for (int x = 0; x < 8; x++) {
    const int ng = P->ngroups > x ? (P->ngroups - x + 7) / 8 : 0;
    glen[x] = ng * nom;
  }
  hipError_t e = hipMemcpyAsync(P->d_qs, P->q.data(), sizeof(Chi0Q) * nq, hipMemcpyHostToDevice, stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(P->d_groups, groups.data(), sizeof(int4) * groups.size(), hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) e = hipMemcpyAsync(P->d_glen, glen, sizeof(glen), hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) {
    delete P;
    return PYGW_E_HIP;
  }
  *out = P;
  return PYGW_OK;
}

// Launch-only (no allocation, copy or memset: capturable in a HIP graph).
int chi0_batch_run(const Chi0BatchPlan* P, int G, hipStream_t stream) {
  using namespace cm;
  if (!P || G < 1) return PYGW_E_ARG;
  const int nom = P->nom;
  for (int q = 0; q < P->nq; q++) {
    const Chi0Q& d = P->q[q];
    float* cs = (float*)d.cs;
    unsigned int* cmax = (unsigned int*)(cs + 2 * d.pb);
    double* wsc = (double*)d.ws;
    unsigned long long* wmax = (unsigned long long*)(wsc + nom);
    const double* kc = P->kcw[q];
    hipLaunchKernelGGL(pygw_zero_u32, dim3(1), dim3(256), 0, stream, cmax, (long)d.pb);
    hipLaunchKernelGGL(pygw_zero_u32, dim3(1), dim3(256), 0, stream, (unsigned int*)wmax, 2L * nom);
    hipLaunchKernelGGL(pygw_colmax, dim3(ceil_div(d.pb, 64), colmax_rows_grid(d.np)), dim3(256), 0, stream, d.M, 0L,
                       d.np, d.pb, 1, cmax);
    hipLaunchKernelGGL(pygw_colscale, dim3(ceil_div(d.pb, 256)), dim3(256), 0, stream, cmax, d.pb, cs);
    hipLaunchKernelGGL(pygw_wmax, dim3(64, nom), dim3(256), 0, stream, kc, d.np, nom, 1, P->coef, wmax);
    hipLaunchKernelGGL(pygw_wscale, dim3(ceil_div(nom, 256)), dim3(256), 0, stream, wmax, nom, wsc);
    hipLaunchKernelGGL(pygw_zero_c128, dim3(2048), dim3(256), 0, stream, d.eps, (long)d.pb * d.pb * nom);
    const long nel = (long)d.npc * 4 * d.pbp;
    int blocks = (int)((nel + 255) / 256);
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(pygw_chi0_presplit, dim3(blocks), dim3(256), 0, stream, d.M, 0L, d.np, d.pb, d.pbp, d.npc, 1,
                       (const float*)cs, (h8*)d.S);
    const long nw_el = (long)nom * d.npc * KC;
    int wblocks = (int)((nw_el + 255) / 256);
    if (wblocks > 65536) wblocks = 65536;
    hipLaunchKernelGGL(pygw_chi0_wprep, dim3(wblocks), dim3(256), 0, stream, kc, d.np, nom, 0, 1, d.npc * KC, P->coef,
                       (const double*)wsc, (float*)d.wf);
  }
  hipLaunchKernelGGL(pygw_zero_u32_small, dim3(1), dim3(64), 0, stream, P->d_counters, 8);
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) ncu = v;
  }
  const int grid = (ncu + 7) / 8 * 8;
  hipLaunchKernelGGL(pygw_chi0_mixed_batch, dim3(grid), dim3(NT), 0, stream, (const Chi0Q*)P->d_qs,
                     (const int4*)P->d_groups, (const int*)P->d_glen, nom, G, P->d_counters);
  for (int q = 0; q < P->nq; q++) {
    const Chi0Q& d = P->q[q];
    hipLaunchKernelGGL(pygw_chi0_mirror, dim3(ceil_div((long)d.pb * d.pb, 256), nom), dim3(256), 0, stream, d.eps,
                       d.pb, nom);
  }
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

void chi0_batch_free(Chi0BatchPlan* P) { delete P; }
