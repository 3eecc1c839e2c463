// C ABI of the Coulomb v^(1/2) basis (Coul_setev, pygw.py:2505-2545) and its
// application to the matrix elements (calc_minm, pygw.py:2666, 2684).
#include "common.h"

size_t coul_setev_ws_bytes(int mbsize);
int coul_setev_launch(const void* vmat, int n0, double evtol, const void* wi0, int iop_coul_x, double* ev_out,
                      double* barcev, void* barcvm, void* vecs, int* matsize, int* immax, int* sweeps, void* ws,
                      size_t ws_size, hipStream_t stream);
int minm_vhalf_launch(const void* mmat, long nrows, int nin, long ldm, const void* barcvm, long ldb, int matsize,
                      void* minm, int out_c64, hipStream_t stream);

namespace {
constexpr int kMaxMb = 16384;
}

extern "C" int pygw_coul_setev_workspace(int mbsize, size_t* bytes) {
  if (mbsize <= 0 || mbsize > kMaxMb || !bytes) return PYGW_E_ARG;
  *bytes = coul_setev_ws_bytes(mbsize);
  return PYGW_OK;
}

extern "C" int pygw_coul_setev(const void* vmat, int mbsize, double evtol, const void* wi0, int iop_coul_x,
                               double* ev, double* barcev, void* barcvm, void* vecs, int* matsize, int* immax,
                               int* n_sweeps, void* workspace, size_t ws_bytes, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (!vmat || !ev || !barcev || !barcvm || !matsize || mbsize <= 0 || mbsize > kMaxMb) return PYGW_E_ARG;
  void* ws = workspace;
  bool own = false;
  if (!ws) {
    ws_bytes = coul_setev_ws_bytes(mbsize);
    HIP_TRY(hipMallocAsync(&ws, ws_bytes, stream));
    own = true;
  }
  int rc = coul_setev_launch(vmat, mbsize, evtol, wi0, iop_coul_x, ev, barcev, barcvm, vecs, matsize, immax, n_sweeps,
                             ws, ws_bytes, stream);
  if (own) {
    hipError_t e = hipFreeAsync(ws, stream);
    if (e != hipSuccess && rc == PYGW_OK) rc = PYGW_E_HIP;
  }
  return rc;
}

extern "C" int pygw_minm_vhalf(const void* mmat, long n_rows, int n_in, long ld_mmat, const void* barcvm,
                               int ld_barcvm, int matsize, int complex64_out, void* minm, void* stream) {
  if (!mmat || !barcvm || !minm || n_rows <= 0 || n_in <= 0 || matsize <= 0 || ld_mmat < n_in ||
      ld_barcvm < n_in || matsize > 65535 * 64 || (complex64_out != 0 && complex64_out != 1))
    return PYGW_E_ARG;
  return minm_vhalf_launch(mmat, n_rows, n_in, ld_mmat, barcvm, ld_barcvm, matsize, minm, complex64_out,
                           (hipStream_t)stream);
}
