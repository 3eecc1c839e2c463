// Shared device/host helpers for the PyGW polarization path on gfx950.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PYGW_OK 0
#define PYGW_E_ARG -10
#define PYGW_E_HIP -11
#define PYGW_E_GEOM -12
#define PYGW_E_NOTPD -13
#define PYGW_E_NOCONV -14
#define PYGW_E_SCRATCH -15

#define HIP_TRY(x)                                  \
  do {                                              \
    hipError_t _e = (x);                            \
    if (_e != hipSuccess) return PYGW_E_HIP;        \
  } while (0)

// Fortran default-real literals widened to real(8).
#define F32C(x) ((double)(float)(x))

static inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// Zero-fill kernel for small device counters (used instead of hipMemsetAsync
// so that callers may capture the launches in a HIP graph).
__global__ void pygw_fill_i32(int* __restrict__ x, int n, int v);
