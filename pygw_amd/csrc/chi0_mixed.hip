// Split-precision chi0 contraction on gfx950 MFMA (v_mfma_f32_32x32x16_f16).
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
//   column side: b = (s_J M[p,J]) (w_p t_w) for a column n = (J, w) of the
//            merged column-frequency index, weighted and split while it is
//            staged into LDS.
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
// Work items.  Rows I come in tiles of 128.  For I tile It the columns are
// the merged list m = 0 .. (PB - 128 It) nom - 1 of (J, w) = (128 It + m/nom,
// m % nom): every frequency of every column J on or right of the tile's
// diagonal block.  One work item = one I tile x 128 consecutive merged
// columns, so the padding is confined to the last I tile and the last column
// chunk of each I tile (at PB 900, 64 frequencies: 1.14x the Hermitian area,
// against 1.46x for 128 x 128 (I, J) tile pairs one frequency at a time).
// The result is eps[I, J, w] for J >= I; pygw_chi0_mirror fills the lower
// triangle.  In the diagonal tiles a wave skips its 32 x 32 tiles that lie
// wholly below the diagonal.
//
// Tiling.  512 threads = 8 waves per block, block tile 128 (merged columns,
// the MFMA rows) x 128 (I, the MFMA columns: the FP64 folds of a half-wave
// touch 512 contiguous bytes of the column-major eps).  The waves split the
// work: waves 0-3 (one per SIMD: a block's waves are dealt to the four SIMDs
// round robin) each compute a 64 x 64 tile, four 32x32 MFMA tiles with
// eight accumulator chains, from the LDS slabs of the current chunk; waves
// 4-7 stage, loading the chunk after next into registers and storing the
// next one into the other LDS buffer (I planes copied, column side weighted
// and split; 2 x 2 x 32 KB).  Each SIMD thus runs one MFMA wave beside one
// staging wave, whose VALU and memory work overlap the MFMAs instead of
// alternating with them at every chunk barrier.  Blocks are dealt to the 8
// XCDs round robin and each XCD takes consecutive work items (same I tile),
// so the I slab they share is read from HBM once per XCD and then served
// from its L2.
//
// Rounds.  All items of one call cost the same, so a launch of n items on B
// CUs (one block per CU) runs floor(n/B) full rounds and a partial one.  The
// r items of the partial round are each cut into S = B/r parts along the
// pairs; a part sums into its own FP64 slot and pygw_chi0_tailfix adds the
// slots of an item in part order (deterministic).
#include "common.h"

#include <algorithm>
#include <type_traits>
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
constexpr int TB = 128;  // tile edge: I rows, merged columns
constexpr int KC = 32;   // pairs per chunk = MFMA K
constexpr int NT = 512;  // threads per block
constexpr int kSlots = 256;  // FP64 part slots for the split items of a launch

// merged columns of I tile It and their 128-wide chunks
__host__ __device__ inline int ncols(int pb, int nom, int It) { return (pb - It * TB) * nom; }
__host__ __device__ inline int ncchunks(int pb, int nom, int It) { return (ncols(pb, nom, It) + TB - 1) / TB; }
// work item t of the It-major list -> (It, Nc)
__device__ inline void item_of(int t, int pb, int nom, int& It, int& Nc) {
  int i = 0;
  for (int n = ncchunks(pb, nom, 0); t >= n; n = ncchunks(pb, nom, ++i)) t -= n;
  It = i;
  Nc = t;
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

// The same split for two values at once, as packed FP16 pairs: hi by one
// packed conversion, x - hi (exact in FP32) by v_fma_mix_f32 with the FP16
// half of hi as an operand, lo by one packed conversion.
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void split2(float x0, float x1, unsigned& hi, unsigned& lo) {
  asm volatile("" : "+v"(x0), "+v"(x1));
  hi = __builtin_bit_cast(unsigned, h2{(_Float16)x0, (_Float16)x1});
  float d0, d1;
  asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(d0) : "v"(hi), "v"(x0));
  asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(d1) : "v"(hi), "v"(x1));
  lo = __builtin_bit_cast(unsigned, h2{(_Float16)d0, (_Float16)d1});
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
// A = the merged columns (rows of D), B = the I side (columns of D).
// S: fragment planes of this slab; M: the same slab as complex64 (column
// side); wt: the slab's scaled weights coef*kcw*t_w in FP32,
// [k][p/8][w][p%8] (zero past N_p), from pygw_chi0_wprep.
// One work item: rows I of tile It against the 128 merged columns
// m = 128 Nc .. 128 Nc + 127 of tile It (see the file header), summed over
// the chunks [ch0, ch1) of the slab (chunk ch = k*npc + c).  With part ==
// nullptr the sum is added into eps[I, J, w]; otherwise it is written into
// the part slot part[row][I - 128 It] (FP64, the first fold stores).
//
// The block's 8 waves split the work: waves 0-3 (one per SIMD) only compute,
// each a 64 x 64 tile (four 32x32 MFMA tiles, eight accumulator chains)
// from the LDS slabs of the current chunk; waves 4-7 only stage, loading the
// chunk after next into registers and storing the next one (I planes copied,
// column side weighted and split) into the other LDS buffer.  A SIMD then
// holds one MFMA wave and one staging wave, whose VALU and memory work run
// beside the MFMAs.  As and Bs are the block's LDS buffers; every thread of
// the 512-thread block calls this with the same arguments, and it ends on a
// block barrier.
__device__ __forceinline__ void chi0_cols(const h8* __restrict__ S, const float2* __restrict__ M, long strideK,
                                          int np, int pb, int pbp, int npc, long ch0, long ch1,
                                          const float* __restrict__ wt, int nom, const float* __restrict__ cs,
                                          const double* __restrict__ ws, int G, int It, int Nc,
                                          double2* __restrict__ eps, double2* __restrict__ part,
                                          h8 (*As)[16][cm::TB], h8 (*Bs)[16][cm::TB]) {
  using namespace cm;
  const int I0 = It * TB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const float* ics = cs + pb;  // 1/cs, exact
  const int mcount = ncols(pb, nom, It);
  const long nch = ch1 - ch0;
  const long last = nch - 1;

  if (wave >= 4) {
    // ---------------- staging waves ----------------
    const int pt = tid - 256;
    // column side: merged column nb, pair groups 2 hf and 2 hf + 1 (8 pairs each)
    const int nb = pt & 127, hf = pt >> 7;
    const int mb = Nc * TB + nb;
    const bool nin = mb < mcount;
    const int jn = nin ? I0 + mb / nom : pb - 1;  // past the end: column PB-1 (finite), zeroed by sJ = 0
    const int wn = mb % nom;
    const float sJ = nin ? cs[jn] : 0.f;
    struct Pre {
      h8 a[8];
      float2 m[16];
      float w[16];
    };
    // Loads go through buffer resources built from chunk-dependent bases
    // (scalar registers) with per-lane byte offsets; out-of-range offsets
    // read as zero: pairs p >= N_p fall past the end of the k-point's M block.
    constexpr int RSRC_FLAGS = 0x00020000;
    const unsigned pb8 = (unsigned)pb * 8u;
    const int npad = npc * KC;
    const unsigned nom4 = (unsigned)nom * 4u;
    const unsigned aoff = (unsigned)((pt >> 7) * pbp + (pt & 127)) * 16u, aq = 2u * (unsigned)pbp * 16u;
    const unsigned moff = (unsigned)(16 * hf) * pb8 + (unsigned)jn * 8u;
    const unsigned woff = ((unsigned)(2 * hf) * (unsigned)nom + (unsigned)wn) * 32u;
    auto fetch = [&](Pre& P, int k, int c) {
      const h8* Sk = S + ((size_t)k * npc + c) * 16 * (size_t)pbp + I0;
      const __amdgpu_buffer_rsrc_t ra =
          __builtin_amdgcn_make_buffer_rsrc((void*)Sk, (short)0, (int)(16u * (unsigned)pbp * 16u), RSRC_FLAGS);
#pragma unroll
      for (int q = 0; q < 8; q++)
        P.a[q] = __builtin_bit_cast(h8, __builtin_amdgcn_raw_buffer_load_b128(ra, aoff + q * aq, 0, 0));
      const __amdgpu_buffer_rsrc_t rm = __builtin_amdgcn_make_buffer_rsrc(
          (void*)(M + (size_t)k * strideK), (short)0, (int)((unsigned)np * pb8), RSRC_FLAGS);
      const unsigned mo = moff + (unsigned)(32 * c) * pb8;
#pragma unroll
      for (int j = 0; j < 16; j++)
        P.m[j] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rm, mo + j * pb8, 0, 0));
      const __amdgpu_buffer_rsrc_t rw = __builtin_amdgcn_make_buffer_rsrc(
          (void*)(wt + ((size_t)k * npad + 32 * c) * nom), (short)0, (int)(32u * nom4), RSRC_FLAGS);
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const unsigned o = woff + (unsigned)h * nom4 * 8u;
        const u4 w0 = __builtin_amdgcn_raw_buffer_load_b128(rw, o, 0, 0);
        const u4 w1 = __builtin_amdgcn_raw_buffer_load_b128(rw, o + 16u, 0, 0);
        P.w[8 * h + 0] = __uint_as_float(w0.x); P.w[8 * h + 1] = __uint_as_float(w0.y);
        P.w[8 * h + 2] = __uint_as_float(w0.z); P.w[8 * h + 3] = __uint_as_float(w0.w);
        P.w[8 * h + 4] = __uint_as_float(w1.x); P.w[8 * h + 5] = __uint_as_float(w1.y);
        P.w[8 * h + 6] = __uint_as_float(w1.z); P.w[8 * h + 7] = __uint_as_float(w1.w);
      }
    };
    auto stage = [&](const Pre& P, int buf) {
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const int idx = pt + 256 * q;
        As[buf][idx >> 7][idx & 127] = P.a[q];
      }
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int gb = 2 * hf + h;
        u4 rh, rl, ih, il;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          unsigned a, b, c, d;
          const int p0 = 8 * h + 2 * j, p1 = p0 + 1;
          split2((P.m[p0].x * sJ) * P.w[p0], (P.m[p1].x * sJ) * P.w[p1], a, b);
          split2((P.m[p0].y * sJ) * P.w[p0], (P.m[p1].y * sJ) * P.w[p1], c, d);
          rh[j] = a;
          rl[j] = b;
          ih[j] = c;
          il[j] = d;
        }
        Bs[buf][4 * gb + 0][nb] = __builtin_bit_cast(h8, rh);
        Bs[buf][4 * gb + 1][nb] = __builtin_bit_cast(h8, rl);
        Bs[buf][4 * gb + 2][nb] = __builtin_bit_cast(h8, ih);
        Bs[buf][4 * gb + 3][nb] = __builtin_bit_cast(h8, il);
      }
    };
    int kf = (int)(ch0 / npc), cf = (int)(ch0 % npc);  // next chunk to fetch
    long nf = 0;                                        // its index, relative to ch0
    auto fetch_next = [&](Pre& P) {
      fetch(P, kf, cf);
      if (nf < last) {
        nf++;
        if (++cf == npc) { cf = 0; kf++; }
      }
    };
    // Pipeline: while the MFMA waves compute chunk ch from buffer ch&1, chunk
    // ch+1 (in registers since the previous step) is stored into the other
    // buffer and chunk ch+2 loads.  Past the end the last chunk is loaded and
    // stored again (harmless: that buffer is not read any more).
    Pre P0, P1;
    fetch_next(P0);
    stage(P0, 0);
    fetch_next(P1);
    __syncthreads();
    for (long ch = 0; ch < nch; ch += 2) {
      fetch_next(P0);  // chunk ch+2
      stage(P1, 1);    // chunk ch+1
      __syncthreads();
      if (ch == last) break;
      fetch_next(P1);  // chunk ch+3
      stage(P0, 0);    // chunk ch+2
      __syncthreads();
    }
    return;
  }

  // ---------------- MFMA waves ----------------
  const int wr = wave >> 1, wc = wave & 1;  // merged-row half, I-column half
  // A 32x32 tile (a, i) of the wave (merged rows 64 wr + 32 a .., columns
  // I0 + 64 wc + 32 i ..) whose columns J all lie below its first I (J < I
  // for every element) is not needed: pygw_chi0_mirror fills eps[I, J] for
  // I > J from eps[J, I].
  int skip[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++) {
    const int m0 = Nc * TB + 64 * wr + 32 * a;
    const int m1 = m0 + 31 < mcount ? m0 + 31 : mcount - 1;
#pragma unroll
    for (int i = 0; i < 2; i++)
      skip[a][i] = __builtin_amdgcn_readfirstlane((m0 >= mcount || m1 / nom < 64 * wc + 32 * i) ? 1 : 0);
  }
  f16v accR[2][2], accI[2][2];
  auto zero = [&]() {
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int v = 0; v < 16; v++) {
          accR[a][i][v] = 0.f;
          accI[a][i][v] = 0.f;
        }
  };
  bool first = true;
  // Per merged row of the item, in LDS once per item: the eps offset
  // (pb w + J) pb (-1 past the end) and the unscale 1/(s_J t_w).  Written by
  // MFMA waves 0-1 before the first barrier of the chunk loop and read in the
  // folds after it.  A fold loads the eps values of 8 accumulator elements
  // at once and then adds and stores them, so the wave waits on one batch of
  // loads instead of on every element.
  __shared__ long rowoff[TB];
  __shared__ double rowf[TB];
  if (tid < TB) {
    const int m = Nc * TB + tid;
    if (m < mcount) {
      const int J = I0 + m / nom, w = m % nom;
      rowoff[tid] = ((long)pb * w + J) * (long)pb;
      rowf[tid] = (double)ics[J] / ws[w];
    } else {
      rowoff[tid] = -1;
      rowf[tid] = 0.0;
    }
  }
  auto flush = [&](auto sk) {
    int Ib = I0 + 64 * wc + (lane & 31), rb = 64 * wr + 4 * (lane >> 5);
    asm volatile("" : "+v"(Ib), "+v"(rb));
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const int I = Ib + 32 * i;
      if (I >= pb) continue;
      const double fi = (double)ics[I];
#pragma unroll
      for (int a = 0; a < 2; a++) {
        if (decltype(sk)::value && skip[a][i]) continue;
#pragma unroll
        for (int h = 0; h < 2; h++) {
          double2 o[8];
          long off[8];
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const int v = 8 * h + u;
            const int row = rb + 32 * a + 8 * (v >> 2) + (v & 3);
            off[u] = rowoff[row] < 0 ? -1 : part ? (long)row * TB + (I - I0) : rowoff[row] + I;
          }
          double2* base = part ? part : eps;
#pragma unroll
          for (int u = 0; u < 8; u++)
            o[u] = (off[u] >= 0 && !(part && first)) ? base[off[u]] : double2{0.0, 0.0};
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const int v = 8 * h + u;
            const int row = rb + 32 * a + 8 * (v >> 2) + (v & 3);
            const double f = fi * rowf[row];
            if (off[u] >= 0)
              base[off[u]] = double2{o[u].x + (double)accR[a][i][v] * f, o[u].y + (double)accI[a][i][v] * f};
          }
        }
      }
    }
    first = false;
  };
  auto products = [&](int buf, auto sk) {
    const int hk = lane >> 5;
    // fragments of K steps 0 and 1; a step-1 fragment is read from LDS as
    // soon as the step-0 fragment it replaces has fed its last MFMA, so the
    // LDS latency of step 1 hides behind the rest of step 0
    h8 A[2][2][4], B[2][2][4];  // [st][i or a][plane]
    auto ldA = [&](int st, int i) {
      const int g = 2 * st + hk, Ic = 64 * wc + 32 * i + (lane & 31);
#pragma unroll
      for (int q = 0; q < 4; q++) A[st][i][q] = As[buf][4 * g + q][Ic];
    };
    auto ldB = [&](int st, int a) {
      const int g = 2 * st + hk, R = 64 * wr + 32 * a + (lane & 31);
#pragma unroll
      for (int q = 0; q < 4; q++) B[st][a][q] = Bs[buf][4 * g + q][R];
    };
    auto tile = [&](int st, int i, int a) {
      if (decltype(sk)::value && skip[a][i]) return;
      const h8 arh = A[st][i][0], arl = A[st][i][1], aih = A[st][i][2], ail = A[st][i][3];
      const h8 brh = B[st][a][0], brl = B[st][a][1], bih = B[st][a][2], bil = B[st][a][3];
      const h8 narh = neg(arh), narl = neg(arl);  // Im = a_i b_r + (-a_r) b_i
      f16v r = accR[a][i], m = accI[a][i];
      r = __builtin_amdgcn_mfma_f32_32x32x16_f16(brh, arh, r, 0, 0, 0);
      m = __builtin_amdgcn_mfma_f32_32x32x16_f16(brh, aih, m, 0, 0, 0);
      r = __builtin_amdgcn_mfma_f32_32x32x16_f16(brl, arh, r, 0, 0, 0);
      m = __builtin_amdgcn_mfma_f32_32x32x16_f16(brl, aih, m, 0, 0, 0);
      r = __builtin_amdgcn_mfma_f32_32x32x16_f16(brh, arl, r, 0, 0, 0);
      m = __builtin_amdgcn_mfma_f32_32x32x16_f16(brh, ail, m, 0, 0, 0);
      r = __builtin_amdgcn_mfma_f32_32x32x16_f16(bih, aih, r, 0, 0, 0);
      m = __builtin_amdgcn_mfma_f32_32x32x16_f16(bih, narh, m, 0, 0, 0);
      r = __builtin_amdgcn_mfma_f32_32x32x16_f16(bil, aih, r, 0, 0, 0);
      m = __builtin_amdgcn_mfma_f32_32x32x16_f16(bil, narh, m, 0, 0, 0);
      r = __builtin_amdgcn_mfma_f32_32x32x16_f16(bih, ail, r, 0, 0, 0);
      m = __builtin_amdgcn_mfma_f32_32x32x16_f16(bih, narl, m, 0, 0, 0);
      accR[a][i] = r;
      accI[a][i] = m;
    };
    // (the scheduling barriers keep the reads where they are written)
    ldA(0, 0);
    ldB(0, 0);
    ldA(0, 1);
    ldB(0, 1);
    tile(0, 0, 0);
    tile(0, 0, 1);
    __builtin_amdgcn_sched_barrier(0);
    ldA(1, 0);   // A(0, 0) done
    __builtin_amdgcn_sched_barrier(0);
    tile(0, 1, 0);
    __builtin_amdgcn_sched_barrier(0);
    ldB(1, 0);   // B(0, 0) done
    __builtin_amdgcn_sched_barrier(0);
    tile(0, 1, 1);
    __builtin_amdgcn_sched_barrier(0);
    ldA(1, 1);   // A(0, 1), B(0, 1) done
    ldB(1, 1);
    __builtin_amdgcn_sched_barrier(0);
    tile(1, 0, 0);
    tile(1, 0, 1);
    tile(1, 1, 0);
    tile(1, 1, 1);
  };

  // the chunk loop, with the tile skips compiled in only where an item has
  // one (so that the products of the others stay one basic block)
  auto run = [&](auto sk) {
    zero();
    __syncthreads();  // chunk 0 staged
    int g = 0;
    for (long ch = 0; ch < nch; ch += 2) {
      products(0, sk);
      if (++g == G || ch == last) {
        flush(sk);
        zero();
        g = 0;
      }
      __syncthreads();
      if (ch == last) break;
      products(1, sk);
      if (++g == G || ch + 1 == last) {
        flush(sk);
        zero();
        g = 0;
      }
      __syncthreads();
    }
  };
  if (skip[0][0] | skip[0][1] | skip[1][0] | skip[1][1])
    run(std::true_type{});
  else
    run(std::false_type{});
}

// One launch over a slab: blocks [0, nfull) take the items of the full
// rounds (XCD-aware: blocks b, b+8, b+16, ... run on one XCD and take
// consecutive items); with sp > 1 each of the remaining items is cut into sp
// parts along the chunks (block nfull + t: item nfull + t/sp, part t%sp,
// slot t), else block b takes item b.
__global__ void __launch_bounds__(512) pygw_chi0_mixed(const h8* __restrict__ S, const float2* __restrict__ M,
                                                       long strideK, int np, int pb, int pbp, int npc, int nkc,
                                                       const float* __restrict__ wt, int nom,
                                                       const float* __restrict__ cs,
                                                       const double* __restrict__ ws, int G, int nfull, int sp,
                                                       double2* __restrict__ eps, double2* __restrict__ part) {
  using namespace cm;
  __shared__ h8 As[2][16][TB];  // [buf][g*4 + plane][I]
  __shared__ h8 Bs[2][16][TB];  // [buf][g*4 + plane][merged column], weighted
  const int b = blockIdx.x;
  const long nch = (long)nkc * npc;
  int item;
  long c0 = 0, c1 = nch;
  double2* slot = nullptr;
  if (b < nfull) {
    const int per = nfull >> 3;
    item = (b & 7) * per + (b >> 3);
  } else if (sp > 1) {
    const int t = b - nfull, s = t % sp;
    item = nfull + t / sp;
    c0 = nch * s / sp;
    c1 = nch * (s + 1) / sp;
    slot = part + (size_t)t * TB * TB;
  } else {
    item = b;
  }
  int It, Nc;
  item_of(item, pb, nom, It, Nc);
  chi0_cols(S, M, strideK, np, pb, pbp, npc, c0, c1, wt, nom, cs, ws, G, It, Nc, eps, slot, As, Bs);
}

// eps[I, J, w] += sum over s of the part slots (t*sp + s) of split item
// nfull + t, in part order.  blockIdx.y = t; one thread per (row, I).
__global__ void __launch_bounds__(256) pygw_chi0_tailfix(const double2* __restrict__ part, int sp, int nfull, int pb,
                                                         int nom, double2* __restrict__ eps) {
  using namespace cm;
  const int t = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = e / TB, i = e % TB;
  int It, Nc;
  item_of(nfull + t, pb, nom, It, Nc);
  const int I = It * TB + i, m = Nc * TB + row;
  if (I >= pb || m >= ncols(pb, nom, It)) return;
  const int J = It * TB + m / nom, w = m % nom;
  if (J < I) return;  // filled by pygw_chi0_mirror (its part slot may be unwritten)
  const double2* p = part + ((size_t)t * sp * TB + row) * TB + i;
  double2 s = p[0];
  for (int q = 1; q < sp; q++) {
    const double2 v = p[(size_t)q * TB * TB];
    s.x += v.x;
    s.y += v.y;
  }
  double2* d = eps + ((size_t)pb * w + J) * (size_t)pb + I;
  const double2 o = *d;
  *d = double2{o.x + s.x, o.y + s.y};
}

// The lower triangle from the upper: eps[I,J] = conj(eps[J,I]) for I > J
// (the contraction computes every eps[I, J] with J >= I, and some with
// J < I inside the diagonal tiles).  One thread per element.
__global__ void pygw_chi0_mirror(double2* __restrict__ eps, int pb, int nom) {
  const long n = (long)pb * pb;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int I = (int)(e % pb), J = (int)(e / pb);  // element (I, J), column-major
  if (I <= J) return;
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

// wt[k][p/8][w][p%8] = (float)(coef * kcw[k0+k][w][p] * ws[w]) for p < N_p,
// 0 up to npad: the column-side weights of one slab, scaled and rounded
// once.  A lane stages 8 consecutive pairs of one frequency (32 contiguous
// bytes) and the lanes of a wave take consecutive frequencies.
__global__ void pygw_chi0_wprep(const double* __restrict__ kcw, int np, int nom, int k0, int nkc, int npad, double coef,
                                const double* __restrict__ ws, float* __restrict__ wt) {
  const long n = (long)nkc * npad * nom;
  const int ng = npad / 8;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int j = (int)(e & 7);
    const long r = e >> 3;  // (k*ng + pg)*nom + w
    const int w = (int)(r % nom);
    const long kg = r / nom;
    const int p = 8 * (int)(kg % ng) + j, k = (int)(kg / ng);
    wt[e] = p < np ? (float)(coef * kcw[((size_t)(k0 + k) * nom + w) * np + p] * ws[w]) : 0.f;
  }
}

namespace {
constexpr size_t kHead = 1 << 16;                // scales region granule, bytes
constexpr size_t kSlabBudget = (size_t)1 << 30;  // default slab budget, bytes
constexpr size_t kPartBytes = (size_t)cm::kSlots * cm::TB * cm::TB * sizeof(double2);

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
int items_total(int pb, int nom) {
  int n = 0;
  for (int It = 0; It * cm::TB < pb; It++) n += cm::ncchunks(pb, nom, It);
  return n;
}
// blocks per round: one per CU, a multiple of 8 (the XCD deal), at most the
// number of part slots
int round_blocks() {
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) ncu = v;
  }
  int b = ncu / 8 * 8;
  if (b < 8) b = 8;
  return b > cm::kSlots ? cm::kSlots : b;
}
}  // namespace

// Workspace for a call over n_k k-points: scales, the part slots of the
// split items, and the fragment planes and weights of up to
// min(n_k, 1 GiB / per-k bytes) k-points (the call walks k in slabs).
size_t chi0_mixed_ws_bytes(int pb, int np, int nom, int nk) {
  const size_t per_k = slab_bytes_per_k(np, pb, nom);
  size_t ks = kSlabBudget / per_k;
  if (ks < 1) ks = 1;
  if (ks > (size_t)nk) ks = (size_t)nk;
  return head_bytes(pb, nom) + kPartBytes + ks * per_k;
}

// M: complex64 [n_kc][np][pb] holding k = k0..k0+n_kc-1; kcw covers all k.
// eps accumulates (beta = 1) or is overwritten (beta = 0).  A workspace
// without room for the part slots (at least head + one k-point) runs the
// last round unsplit.
// Optional timing of the contraction kernel alone (pygw_chi0_launch_timer):
// events recorded on the launch stream around every pygw_chi0_mixed launch,
// their elapsed times summed when read.  For measurement runs, not for code
// under HIP graph capture.
namespace cm {
struct LaunchTimer {
  bool on = false;
  std::vector<hipEvent_t> ev;   // start, stop pairs
  size_t used = 0;
};
static LaunchTimer g_timer;
}  // namespace cm

extern "C" int pygw_chi0_launch_timer(int max_launches) {
  using namespace cm;
  for (hipEvent_t e : g_timer.ev) (void)hipEventDestroy(e);
  g_timer.ev.clear();
  g_timer.used = 0;
  g_timer.on = max_launches > 0;
  for (int i = 0; i < 2 * max_launches; i++) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) {
      g_timer.on = false;
      return PYGW_E_HIP;
    }
    g_timer.ev.push_back(e);
  }
  return PYGW_OK;
}

// total kernel time (ms) and launch count since the timer was armed (or last
// read); synchronises the recorded events
extern "C" int pygw_chi0_launch_time(double* ms, int* launches) {
  using namespace cm;
  if (!ms || !launches) return PYGW_E_ARG;
  double t = 0.0;
  for (size_t i = 0; i + 1 < 2 * g_timer.used + 1 && i < g_timer.ev.size(); i += 2) {
    if (hipEventSynchronize(g_timer.ev[i + 1]) != hipSuccess) return PYGW_E_HIP;
    float x = 0.f;
    if (hipEventElapsedTime(&x, g_timer.ev[i], g_timer.ev[i + 1]) != hipSuccess) return PYGW_E_HIP;
    t += x;
  }
  *ms = t;
  *launches = (int)g_timer.used;
  g_timer.used = 0;
  return PYGW_OK;
}

int chi0_mixed_launch(const void* M, long strideK, int np, int pb, const double* kcw, int nom, int k0, int nkc,
                      double coef, int G, int beta, void* ws, size_t ws_bytes, void* eps, hipStream_t stream) {
  using namespace cm;
  const size_t head = head_bytes(pb, nom);
  const size_t per_k = slab_bytes_per_k(np, pb, nom);
  if (ws_bytes < head + per_k || G < 1) return PYGW_E_ARG;
  if ((size_t)(np + KC) * pb * 8 >= ((size_t)1 << 31)) return PYGW_E_ARG;  // 32-bit buffer offsets per k
  if ((long)pb * nom >= (1L << 30)) return PYGW_E_ARG;                     // 32-bit merged column index
  const bool parts = ws_bytes >= head + kPartBytes + per_k;
  const size_t body = ws_bytes - head - (parts ? kPartBytes : 0);
  size_t ks = body / per_k;
  if (ks > (size_t)nkc) ks = (size_t)nkc;
  float* cs = (float*)ws;  // [pb] scales, then [pb] their inverses
  unsigned int* cmax = (unsigned int*)(cs + 2 * pb);
  double* wsc = (double*)(((uintptr_t)(cmax + pb) + 15) & ~(uintptr_t)15);
  unsigned long long* wmax = (unsigned long long*)(wsc + nom);
  double2* part = parts ? (double2*)((char*)ws + head) : nullptr;
  const int T = ceil_div(pb, TB);
  const int pbp = T * TB;
  const int npc = ceil_div(np, KC);
  h8* S = (h8*)((char*)ws + head + (parts ? kPartBytes : 0));
  float* wt = (float*)((char*)S + ks * (size_t)npc * 16 * pbp * sizeof(h8));
  const float2* Mf = (const float2*)M;

  hipLaunchKernelGGL(pygw_zero_u32, dim3(1), dim3(256), 0, stream, cmax, (long)pb);
  hipLaunchKernelGGL(pygw_zero_u32, dim3(1), dim3(256), 0, stream, (unsigned int*)wmax, 2L * nom);
  hipLaunchKernelGGL(pygw_colmax, dim3(ceil_div(pb, 64), colmax_rows_grid((long)nkc * np)), dim3(256), 0, stream, Mf,
                     strideK, np, pb, nkc, cmax);
  hipLaunchKernelGGL(pygw_colscale, dim3(ceil_div(pb, 256)), dim3(256), 0, stream, cmax, pb, cs);
  hipLaunchKernelGGL(pygw_wmax, dim3(64, nom), dim3(256), 0, stream, kcw + (size_t)k0 * nom * np, np, nom, nkc, coef,
                     wmax);
  hipLaunchKernelGGL(pygw_wscale, dim3(ceil_div(nom, 256)), dim3(256), 0, stream, wmax, nom, wsc);
  const long ntot = (long)pb * pb * nom;
  if (!beta) hipLaunchKernelGGL(pygw_zero_c128, dim3(2048), dim3(256), 0, stream, (double2*)eps, ntot);

  const int nitems = items_total(pb, nom);
  const int B = round_blocks();
  const int nfull = nitems / B * B;
  const int rest = nitems - nfull;
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
                       coef, (const double*)wsc, wt);
    // parts per split item: fill the round, at most one per chunk
    int sp = (parts && rest > 0) ? B / rest : 1;
    if ((long)sp > (long)n * npc) sp = (int)((long)n * npc);
    if (sp < 2) sp = 1;
    const int grid = sp > 1 ? nfull + rest * sp : nitems;
    const bool timed = g_timer.on && 2 * (g_timer.used + 1) <= g_timer.ev.size();
    if (timed) (void)hipEventRecord(g_timer.ev[2 * g_timer.used], stream);
    hipLaunchKernelGGL(pygw_chi0_mixed, dim3(grid), dim3(NT), 0, stream, (const h8*)S, Ms, strideK, np, pb, pbp, npc,
                       n, (const float*)wt, nom, cs, wsc, G, nfull, sp, (double2*)eps, part);
    if (timed) (void)hipEventRecord(g_timer.ev[2 * g_timer.used++ + 1], stream);
    if (sp > 1)
      hipLaunchKernelGGL(pygw_chi0_tailfix, dim3(TB * TB / 256, rest), dim3(256), 0, stream, (const double2*)part, sp,
                         nfull, pb, nom, (double2*)eps);
  }
  hipLaunchKernelGGL(pygw_chi0_mirror, dim3(ceil_div((long)pb * pb, 256), nom), dim3(256), 0, stream,
                     (double2*)eps, pb, nom);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}
// ---------------------------------------------------------------------------
// Batched contraction over many q-points in one launch.  Each q
// has its own PB and pair count (all pairs of all k concatenated: only the
// (k, p) list matters to the sum), its own operands, scales and eps.  Work
// item = (q, I tile, merged-column chunk).  The items, in order of
// decreasing cost (the q's chunk count) and then I tile, are cut into
// segments of 32 (one per CU of an XCD) dealt to the 8 XCDs round robin;
// one block per item, block b taking item b >> 3 of XCD b & 7's list, so
// blocks that run together on an XCD share an I slab in its L2.
// ---------------------------------------------------------------------------
struct Chi0Q {
  const h8* S;        // fragment planes [npc][16][pbp]
  const float2* M;    // complex64 [K][PB]
  const float* wt;    // FP32 scaled weights [npc*4][nom][8]
  const float* cs;    // [2*PB] column scales and inverses
  const double* ws;   // [nom] weight scales
  double2* eps;       // Fortran [PB, PB, nom]
  int pb, pbp, np, npc;
};

__global__ void __launch_bounds__(512) pygw_chi0_mixed_batch(const Chi0Q* __restrict__ qs,
                                                             const int4* __restrict__ items,
                                                             const int* __restrict__ xoff, int nom, int G) {
  using namespace cm;
  __shared__ h8 As[2][16][TB];
  __shared__ h8 Bs[2][16][TB];
  // block b runs on XCD b % 8 (workgroups are dealt to the XCDs round
  // robin, and each XCD starts its workgroups in order) and takes item
  // b / 8 of that XCD's list
  const int x = blockIdx.x & 7, i = blockIdx.x >> 3;
  const int o = xoff[x];
  if (i >= xoff[x + 1] - o) return;
  const int4 it = items[o + i];  // (q, It, Nc)
  const Chi0Q d = qs[it.x];
  chi0_cols(d.S, d.M, 0, d.np, d.pb, d.pbp, d.npc, 0, d.npc, d.wt, nom, d.cs, d.ws, G, it.y, it.z, d.eps,
            nullptr, As, Bs);
}

// Host-side plan of a batch (opaque to callers): per-q sizes and pointers,
// device tables in the caller's workspace.
struct Chi0BatchPlan {
  int nq, nom;
  double coef;
  std::vector<Chi0Q> q;
  std::vector<const double*> kcw;
  Chi0Q* d_qs;
  int4* d_items;
  int* d_xoff;
  int nitems, maxlen;
  char* ws;
};

namespace {
constexpr int kSegment = 32;  // items per segment dealt to one XCD

size_t batch_q_bytes(int pb, int np, int nom) {
  const size_t npc = (size_t)(np + cm::KC - 1) / cm::KC;
  const size_t pbp = (size_t)(pb + cm::TB - 1) / cm::TB * cm::TB;
  return head_bytes(pb, nom) + npc * 16 * pbp * sizeof(h8) +
         ((size_t)nom * npc * cm::KC * sizeof(float) + 255) / 256 * 256;
}
size_t batch_tables_bytes(int nq, const int* pb, int nom) {
  size_t ni = 0;
  for (int q = 0; q < nq; q++) ni += (size_t)items_total(pb[q], nom);
  const size_t a = ((size_t)nq * sizeof(Chi0Q) + 255) / 256 * 256;
  const size_t b = (ni * sizeof(int4) + 255) / 256 * 256;
  return a + b + 512;
}
}  // namespace

size_t chi0_batch_ws_bytes(int nq, const int* pb, const int* np, int nom) {
  size_t t = batch_tables_bytes(nq, pb, nom);
  for (int q = 0; q < nq; q++) t += batch_q_bytes(pb[q], np[q], nom);
  return t;
}

// Builds the plan and uploads its tables (synchronously on `stream`).
int chi0_batch_plan(int nq, const void* const* M, const int* pb, const int* np, const double* const* kcw, int nom,
                    double coef, void* const* eps, void* ws, size_t ws_bytes, hipStream_t stream,
                    Chi0BatchPlan** out) {
  using namespace cm;
  if (ws_bytes < chi0_batch_ws_bytes(nq, pb, np, nom)) return PYGW_E_ARG;
  for (int q = 0; q < nq; q++)
    if ((long)pb[q] * nom >= (1L << 30)) return PYGW_E_ARG;  // 32-bit merged column index
  Chi0BatchPlan* P = new Chi0BatchPlan();
  P->nq = nq;
  P->nom = nom;
  P->coef = coef;
  P->ws = (char*)ws;
  size_t off = batch_tables_bytes(nq, pb, nom);
  std::vector<int4> all;
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
    d.wt = (const float*)(base + head + (size_t)d.npc * 16 * d.pbp * sizeof(h8));
    d.M = (const float2*)M[q];
    d.eps = (double2*)eps[q];
    P->q.push_back(d);
    P->kcw.push_back(kcw[q]);
    off += batch_q_bytes(pb[q], np[q], nom);
    for (int It = 0; It < T; It++)
      for (int c = 0; c < ncchunks(pb[q], nom, It); c++) all.push_back(int4{q, It, c, 0});
  }
  // longest first: an item costs ~ its q's chunk count, so dealing them in
  // decreasing cost leaves the short ones for the tail of each XCD
  std::stable_sort(all.begin(), all.end(),
                   [&](const int4& a, const int4& b) { return P->q[a.x].npc > P->q[b.x].npc; });
  P->nitems = (int)all.size();
  std::vector<int4> items;
  items.reserve(all.size());
  int xoff[9];
  for (int x = 0; x < 8; x++) {
    xoff[x] = (int)items.size();
    for (size_t s = (size_t)x * kSegment; s < all.size(); s += 8 * kSegment)
      for (size_t i = s; i < s + kSegment && i < all.size(); i++) items.push_back(all[i]);
  }
  xoff[8] = (int)items.size();
  P->maxlen = 0;
  for (int x = 0; x < 8; x++) P->maxlen = std::max(P->maxlen, xoff[x + 1] - xoff[x]);
  const size_t a = ((size_t)nq * sizeof(Chi0Q) + 255) / 256 * 256;
  const size_t b = ((size_t)P->nitems * sizeof(int4) + 255) / 256 * 256;
  P->d_qs = (Chi0Q*)ws;
  P->d_items = (int4*)((char*)ws + a);
  P->d_xoff = (int*)((char*)ws + a + b);
  hipError_t e = hipMemcpyAsync(P->d_qs, P->q.data(), sizeof(Chi0Q) * nq, hipMemcpyHostToDevice, stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(P->d_items, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) e = hipMemcpyAsync(P->d_xoff, xoff, sizeof(xoff), hipMemcpyHostToDevice, stream);
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
                       (const double*)wsc, (float*)d.wt);
  }
  hipLaunchKernelGGL(pygw_chi0_mixed_batch, dim3(8 * P->maxlen), dim3(NT), 0, stream, (const Chi0Q*)P->d_qs,
                     (const int4*)P->d_items, (const int*)P->d_xoff, nom, G);
  for (int q = 0; q < P->nq; q++) {
    const Chi0Q& d = P->q[q];
    hipLaunchKernelGGL(pygw_chi0_mirror, dim3(ceil_div((long)d.pb * d.pb, 256), nom), dim3(256), 0, stream, d.eps,
                       d.pb, nom);
  }
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

void chi0_batch_free(Chi0BatchPlan* P) { delete P; }
