// Seeded synthetic matrix elements on the device (benchmark input generator;
// no counterpart in PyGW, whose M comes from calc_minm, pygw.py:2547-2721).
//
//   M[k][p][I] = g(seed, (k*n_pairs + p)*pb + I) * exp(-|e_n(k) - e_m(k-q)|/20 eV)
//                / sqrt(pb) * vhalf[I],     p = i_o*n_e + i_e,
// n = row_band0 + i_o, m = col_band0 + i_e (0-based band indices): the
// polarization pairs use row_band0 = 0, col_band0 = ncg + numin; the
// self-energy rows use the GW bands against all bands.
// g = complex Gaussian with Re, Im ~ N(0, 1/2): SplitMix64 of the counter,
// Box-Muller.  pygw_amd/synth.py holds the host twin of the same draws.
#include "common.h"

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  uint64_t z = x;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// enk: Fortran (nbnd, nk); out complex128 (cplx=1) or complex64 (cplx=0)
__global__ void __launch_bounds__(256) pygw_synth_minm_k(const double* __restrict__ enk, int nbnd, const int* __restrict__ kqid,
                                                         int row0, int col0, int no, int ne, int pb,
                                                         const double* __restrict__ vhalf, uint64_t seed, int k0, int nkc,
                                                         int cplx, void* __restrict__ out) {
  const long np = (long)no * ne;
  const long per_k = np * pb;
  const long n = per_k * nkc;
  const double h2ev = 27.2113961;
  const double rs = 1.0 / sqrt((double)pb);
  const uint64_t base = seed * 0xD1B54A32D192ED03ull;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int kl = (int)(e / per_k);
    const long r = e % per_k;
    const int p = (int)(r / pb), I = (int)(r % pb);
    const int k = k0 + kl;
    const int io = p / ne, ie = p % ne;
    const double eo = enk[(row0 + io) + (size_t)nbnd * k];
    const double ee = enk[(col0 + ie) + (size_t)nbnd * kqid[k]];
    const double sc = exp(-fabs(eo - ee) * h2ev / 20.0) * rs * vhalf[I];
    const uint64_t idx = (uint64_t)k * (uint64_t)per_k + (uint64_t)r;
    const uint64_t a = splitmix64(base ^ (idx * 2ull));
    const uint64_t b = splitmix64(base ^ (idx * 2ull + 1ull));
    const double u1 = ((double)(a >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    const double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
    const double rr = sqrt(-log(u1));
    double s, c;
    sincos(2.0 * 3.141592653589793 * u2, &s, &c);
    const double re = rr * c * sc, im = rr * s * sc;
    if (cplx) ((double2*)out)[e] = double2{re, im};
    else ((float2*)out)[e] = float2{(float)re, (float)im};
  }
}

extern "C" int pygw_synth_minm(const double* enk, int n_bnd, const int* kqid, int row_band0, int col_band0, int n_o,
                               int n_e, int pb, const double* vhalf, unsigned long long seed, int k0, int n_kc,
                               int complex128, void* out, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (!enk || !kqid || !vhalf || !out || n_o <= 0 || n_e <= 0 || pb <= 0 || n_kc <= 0 || k0 < 0) return PYGW_E_ARG;
  if (row_band0 < 0 || col_band0 < 0 || col_band0 + n_e > n_bnd || row_band0 + n_o > n_bnd) return PYGW_E_ARG;
  const long n = (long)n_o * n_e * pb * n_kc;
  int blocks = (int)((n + 255) / 256);
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(pygw_synth_minm_k, dim3(blocks), dim3(256), 0, stream, enk, n_bnd, kqid, row_band0, col_band0, n_o, n_e,
                     pb, vhalf, (uint64_t)seed, k0, n_kc, complex128, out);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}

// Rows of the same draws for a list of (k, p) pairs: out[j][I] = M[k_j][p_j][I]
// of the generator above (idx = (k*n_o*n_e + p)*pb + I), used to build ragged
// per-k pair windows.
__global__ void __launch_bounds__(256) pygw_synth_minm_pairs_k(const double* __restrict__ enk, int nbnd,
                                                              const int* __restrict__ kqid, int row0, int col0,
                                                              int no, int ne, int pb,
                                                              const double* __restrict__ vhalf, uint64_t seed,
                                                              const int2* __restrict__ pairs, long npairs, int cplx,
                                                              void* __restrict__ out) {
  const long np = (long)no * ne;
  const long n = npairs * pb;
  const double h2ev = 27.2113961;
  const double rs = 1.0 / sqrt((double)pb);
  const uint64_t base = seed * 0xD1B54A32D192ED03ull;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long j = e / pb;
    const int I = (int)(e % pb);
    const int2 kp = pairs[j];
    const int k = kp.x, p = kp.y;
    const int io = p / ne, ie = p % ne;
    const double eo = enk[(row0 + io) + (size_t)nbnd * k];
    const double ee = enk[(col0 + ie) + (size_t)nbnd * kqid[k]];
    const double sc = exp(-fabs(eo - ee) * h2ev / 20.0) * rs * vhalf[I];
    const uint64_t idx = ((uint64_t)k * (uint64_t)np + (uint64_t)p) * (uint64_t)pb + (uint64_t)I;
    const uint64_t a = splitmix64(base ^ (idx * 2ull));
    const uint64_t b = splitmix64(base ^ (idx * 2ull + 1ull));
    const double u1 = ((double)(a >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    const double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
    const double rr = sqrt(-log(u1));
    double sn, cn;
    sincos(2.0 * 3.141592653589793 * u2, &sn, &cn);
    const double re = rr * cn * sc, im = rr * sn * sc;
    if (cplx) ((double2*)out)[e] = double2{re, im};
    else ((float2*)out)[e] = float2{(float)re, (float)im};
  }
}

extern "C" int pygw_synth_minm_pairs(const double* enk, int n_bnd, const int* kqid, int row_band0, int col_band0,
                                     int n_o, int n_e, int pb, const double* vhalf, unsigned long long seed,
                                     const int* pairs, long n_pairs, int complex128, void* out, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (!enk || !kqid || !vhalf || !pairs || !out || n_o <= 0 || n_e <= 0 || pb <= 0 || n_pairs <= 0) return PYGW_E_ARG;
  if (row_band0 < 0 || col_band0 < 0 || col_band0 + n_e > n_bnd || row_band0 + n_o > n_bnd) return PYGW_E_ARG;
  const long n = n_pairs * pb;
  int blocks = (int)((n + 255) / 256);
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(pygw_synth_minm_pairs_k, dim3(blocks), dim3(256), 0, stream, enk, n_bnd, kqid, row_band0,
                     col_band0, n_o, n_e, pb, vhalf, (uint64_t)seed, (const int2*)pairs, n_pairs, complex128, out);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}
