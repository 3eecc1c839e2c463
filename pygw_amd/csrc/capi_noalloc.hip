// C-ABI: the dielectric inversion with no allocation of any kind
// (pygw_eps_inverse_noalloc, include/pygw_hip.h).
//
// pygw_eps_inverse needs a workspace (pivot-block inverses, the two panels of
// every block step, per-frequency status words).  Callers that cannot
// allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid it) get the same kernels on a fixed file-scope device
// array instead.  Each device holds its own copy of the array, from the
// library's load until the process ends; its device address is looked up once
// per device (hipGetSymbolAddress, with that device current) and kept.  Every
// call on a device shares that device's array and leaves its contents behind:
// calls that overlap on one device (two streams, two threads) race.
#include "common.h"
#include "../../include/pygw_hip.h"

#include <mutex>

int eps_inverse_launch(void* eps, int n, int nom, void* ws, int* dinfo, hipStream_t stream);
size_t eps_inverse_ws_bytes(int n, int nom);

namespace {
constexpr size_t kScratchBytes = PYGW_NOALLOC_SCRATCH_BYTES;   // PB 900 x 68 frequencies, PB 1200 x 51
constexpr int kMaxDevices = 64;

// dinfo[nom] = OR of the per-frequency status words (one block)
__global__ void __launch_bounds__(256) pygw_na_or_i32(int* __restrict__ d, int n) {
  __shared__ int s[256];
  int v = 0;
  for (int i = threadIdx.x; i < n; i += 256) v |= d[i];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] |= s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) d[n] = s[0];
}

std::mutex g_mu;
void* g_ptr[kMaxDevices];   // device address of `scratch` on each device (null: not looked up yet)
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The device address of this device's scratch array (looked up on the first
// call on the device, with it current).  nullptr on a HIP error.
static void* scratch_ptr() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    g_ptr[dev] = p;
  }
  return g_ptr[dev];
}

extern "C" int pygw_eps_inverse_noalloc(void* eps, int pb, int n_om, int* info, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (!eps || pb <= 0 || n_om <= 0 || n_om > 65535) return PYGW_E_ARG;
  const size_t need = eps_inverse_ws_bytes(pb, n_om);
  if (need > kScratchBytes) return PYGW_E_SCRATCH;
  void* ws = scratch_ptr();
  if (!ws) return PYGW_E_HIP;
  // the status words sit at the end of the used part, one more word for their OR
  int* dinfo = (int*)((char*)ws + need - 256 - sizeof(int) * (size_t)n_om);
  hipLaunchKernelGGL(pygw_fill_i32, dim3(ceil_div(n_om + 1, 256)), dim3(256), 0, stream, dinfo, n_om + 1, 0);
  int rc = eps_inverse_launch(eps, pb, n_om, ws, dinfo, stream);
  if (rc == PYGW_OK && info) {
    hipLaunchKernelGGL(pygw_na_or_i32, dim3(1), dim3(256), 0, stream, dinfo, n_om);
    int bad = 0;
    if (hipMemcpyAsync(&bad, dinfo + n_om, sizeof(int), hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
      return PYGW_E_HIP;
    info[0] = bad;
    if (bad) rc = PYGW_E_NOTPD;
  }
  return rc;
}
