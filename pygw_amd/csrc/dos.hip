// Tetrahedron integrated density of states and density of states at a list
// of energies (idos / dostet, for_tetrahedra.f90:84-141, with the
// one-tetrahedron forms intdos1t / dos1t of Bloechl et al., PRB 49, 16223
// (1994), for_tetrahedra2.f90:108-212).  Used for the quasiparticle Fermi
// level of the GW0 loop (calc_Fermi, mcommon.py:34-66).
#include "common.h"

namespace dos {
// ascending insertion sort of the four corner energies
__device__ __forceinline__ void sort4(double* e) {
  for (int i = 1; i < 4; i++) {
    const double x = e[i];
    int j = i - 1;
    while (j >= 0 && e[j] > x) {
      e[j + 1] = e[j];
      j--;
    }
    e[j + 1] = x;
  }
}

// number of corners at or below ei (4 when the top corner is)
__device__ __forceinline__ int below(const double* et, double ei) {
  if (et[3] <= ei) return 4;
  int n = 0;
  for (int i = 0; i < 4; i++) n += et[i] <= ei;
  return n;
}

__device__ __forceinline__ double intdos1t(const double* et, double ei) {
  switch (below(et, ei)) {
    case 1: {
      const double e21 = et[1] - et[0], e31 = et[2] - et[0], e41 = et[3] - et[0], x = ei - et[0];
      return x * x * x / (e21 * e31 * e41);
    }
    case 2: {
      const double e21 = et[1] - et[0], e31 = et[2] - et[0], e41 = et[3] - et[0];
      const double e32 = et[2] - et[1], e42 = et[3] - et[1], x = ei - et[1];
      const double br = e21 * e21 + 3.0 * x * e21 + 3.0 * x * x - (e31 + e42) * x * x * x / (e32 * e42);
      return br / (e31 * e41);
    }
    case 3: {
      const double e43 = et[3] - et[2], e42 = et[3] - et[1], e41 = et[3] - et[0], x = et[3] - ei;
      return 1.0 - x * x * x / (e41 * e42 * e43);
    }
    case 4:
      return 1.0;
    default:
      return 0.0;
  }
}

__device__ __forceinline__ double dos1t(const double* et, double ei) {
  switch (below(et, ei)) {
    case 1: {
      const double e21 = et[1] - et[0], e31 = et[2] - et[0], e41 = et[3] - et[0], x = ei - et[0];
      return 3.0 * x * x / (e21 * e31 * e41);
    }
    case 2: {
      const double e21 = et[1] - et[0], e31 = et[2] - et[0], e41 = et[3] - et[0];
      const double e32 = et[2] - et[1], e42 = et[3] - et[1], x = ei - et[1];
      const double br = e21 + 2.0 * x - (e31 + e42) * x * x / (e32 * e42);
      return 3.0 * br / (e31 * e41);
    }
    case 3: {
      const double e43 = et[3] - et[2], e42 = et[3] - et[1], e41 = et[3] - et[0], x = et[3] - ei;
      return 3.0 * x * x / (e41 * e42 * e43);
    }
    default:
      return 0.0;
  }
}
}  // namespace dos

// idos[e] = sum_{band, tet} wtet * intdos1t, dos[e] likewise with dos1t, at
// energies[e].  ebnd [n_k][n_band] (row-major), atet [n_tet][4] (k indices).
// One block per energy; the threads stride over (band, tetrahedron) and the
// partial sums are reduced in a fixed order.
__global__ void __launch_bounds__(256) pygw_tetra_idos(const double* __restrict__ ebnd, int nband,
                                                       const int* __restrict__ atet,
                                                       const double* __restrict__ wtet, int ntet,
                                                       const double* __restrict__ energies,
                                                       double* __restrict__ idos, double* __restrict__ dos_out) {
  __shared__ double ri[4], rd[4];
  const double E = energies[blockIdx.x];
  double si = 0.0, sd = 0.0;
  const long n = (long)nband * ntet;
  for (long t = threadIdx.x; t < n; t += blockDim.x) {
    const int ib = (int)(t / ntet), it = (int)(t % ntet);
    double e[4];
#pragma unroll
    for (int i = 0; i < 4; i++) e[i] = ebnd[(size_t)atet[4 * (size_t)it + i] * nband + ib];
    dos::sort4(e);
    si += dos::intdos1t(e, E) * wtet[it];
    sd += dos::dos1t(e, E) * wtet[it];
  }
  for (int off = 32; off > 0; off >>= 1) {
    si += __shfl_xor(si, off, 64);
    sd += __shfl_xor(sd, off, 64);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    ri[w] = si;
    rd[w] = sd;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    idos[blockIdx.x] = (ri[0] + ri[1]) + (ri[2] + ri[3]);
    dos_out[blockIdx.x] = (rd[0] + rd[1]) + (rd[2] + rd[3]);
  }
}

int tetra_idos_launch(const double* ebnd, int nband, const int* atet, const double* wtet, int ntet,
                      const double* energies, int ne, double* idos, double* dos_out, hipStream_t stream) {
  hipLaunchKernelGGL(pygw_tetra_idos, dim3(ne), dim3(256), 0, stream, ebnd, nband, atet, wtet, ntet, energies, idos,
                     dos_out);
  return hipGetLastError() == hipSuccess ? PYGW_OK : PYGW_E_HIP;
}
