// C ABI of the Coulomb v^(1/2) basis (Coul_setev, pygw.py:2505-2545) and its
// application to the matrix elements (calc_minm, pygw.py:2666, 2684).
#include "common.h"

size_t coul_setev_ws_bytes(int mbsize);
int coul_setev_launch(const void* vmat, int n0, double evtol, const void* wi0, int iop_coul_x, double* ev_out,
                      double* barcev, void* barcvm, void* vecs, int* matsize, int* immax, int* sweeps, void* ws,
                      size_t ws_size, hipStream_t stream);
int minm_vhalf_launch(const void* mmat, long nrows, int nin, long ldm, const void* barcvm, long ldb, int matsize,
                      void* minm, int out_c64, hipStream_t stream);

size_t minm_is_ws_bytes(int ngq, int nvi, int nvj, int ndim);
size_t minm_is_ws_bytes_dense(int ngq, int ngq_barc, int nvi, int nvj, int ndim);
size_t minm_is_ws_bytes_rho(int ngq, int ngq_barc, int nvi, int nvj, int ndim, int mdim);
bool minm_is_rho_cheaper(int ngq, int ngq_barc, int nvj, int ndim, int mdim);
int minm_is_launch(int nstart, int nend, int mstart, int mend, const void* Aeigk, int ldk, const void* Aeigq,
                   int ldq, const int* umk, const void* mpwipw, int ngq, int ngq_barc, int nvi, int nvj,
                   const int* indgk_i, const int* indgk_j, const int* inv, int inv_size, const int* gindex,
                   int ng, const int* i_g0, int n1, int n2, int n3, double vol, void* mmat, long ldo, int col0, void* ws,
                   size_t ws_size, hipStream_t stream);

size_t pw_basis_ws_bytes(int ngq, int ngq_barc);
int pw_basis_launch(const int* indgq, int ngq, int ngq_barc, const void* ipwint, int ng, const int* gindex,
                    const int* i_g0, int n1, int n2, int n3, void* mpwipw, void* ws, size_t ws_size,
                    hipStream_t stream);
int coul_pw_block_launch(const void* mpwipw, int ngq, int ngq_barc, const double* gqlen, int q_is_zero, void* vmat,
                         long ld, int row0, hipStream_t stream);

int mixed_coulomb_launch(const double* vq, int iat, int q_vq, const double* jlam, int ngs, const int* gindex,
                         const int* indgq, const double* gqlen0, const int* G_unique, int ngq_barc, const int* big_l,
                         int nmix, const double* k2cartes, const double* rotloc, const double* trotij, const int* mult,
                         const double* pos, int nat, double vol, void* out, long ld, int row0, void* sph_ws,
                         hipStream_t stream);
size_t mixed_coulomb_ws_bytes(int max_mult, int ngq_barc);
int coul_mixed_block_launch(const void* mpwipw, int ngq, int ngq_barc, const void* vmatit, int loct, void* vmat,
                            long ld, hipStream_t stream);

namespace {
constexpr int kMaxMb = 16384;
constexpr int kMaxBarc = 4096;                  // gathered form of pygw_minm_is (LDS row)
constexpr size_t kDenseCap = (size_t)4 << 30;    // largest workspace pygw_minm_is allocates itself
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

// the workspace of the form pygw_minm_is prefers for these sizes: the pair
// density when it costs fewer products, else the dense first product
static size_t minm_is_preferred_bytes(int ngq, int ngq_barc, int nvi, int nvj, int ndim, int mdim) {
  return minm_is_rho_cheaper(ngq, ngq_barc, nvj, ndim, mdim)
             ? minm_is_ws_bytes_rho(ngq, ngq_barc, nvi, nvj, ndim, mdim)
             : minm_is_ws_bytes_dense(ngq, ngq_barc, nvi, nvj, ndim);
}

extern "C" int pygw_minm_is_workspace(int ngq, int ngq_barc, int nvi, int nvj, int ndim, int mdim, size_t* bytes) {
  if (ngq <= 0 || ngq_barc <= 0 || nvi <= 0 || nvj <= 0 || ndim <= 0 || mdim <= 0 || !bytes) return PYGW_E_ARG;
  *bytes = minm_is_preferred_bytes(ngq, ngq_barc, nvi, nvj, ndim, mdim);
  return PYGW_OK;
}

extern "C" int pygw_minm_is(int nstart, int nend, int mstart, int mend, const void* Aeigk, int nbmax_i, int ngi,
                            const void* Aeigq, int nbmax_j, int ngj, const int* iumklap, const void* mpwipw, int ngq,
                            int ngq_barc, int nvi, int nvj, const int* indgk_i, const int* indgk_j,
                            const int* indgq_inverse, int inv_size, const int* gindex, int ng, const int* i_g0, int n1,
                            int n2, int n3, double vol, void* mmat, long ld_mmat, int col0, void* workspace,
                            size_t ws_bytes, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (!Aeigk || !Aeigq || !iumklap || !mpwipw || !indgk_i || !indgk_j || !indgq_inverse || !gindex || !i_g0 ||
      !mmat)
    return PYGW_E_ARG;
  if (nstart < 0 || nend <= nstart || nend > nbmax_i || mstart < 0 || mend <= mstart || mend > nbmax_j ||
      nvi <= 0 || nvi > ngi || nvj <= 0 || nvj > ngj || ngq <= 0 || ngq > 65535 || ngq_barc <= 0 ||
      inv_size <= 0 || ng <= 0 || n1 < 0 || n2 < 0 || n3 < 0 || !(vol > 0.0) || col0 < 0 ||
      ld_mmat < (long)col0 + ngq)
    return PYGW_E_ARG;
  const int ndim = nend - nstart;
  void* ws = workspace;
  bool own = false;
  if (!ws) {
    // the preferred form's buffers (pair density or dense P) when they are
    // at most kDenseCap and can be had, else the gathered form's (kmap + T)
    ws_bytes = minm_is_preferred_bytes(ngq, ngq_barc, nvi, nvj, ndim, mend - mstart);
    if (ws_bytes > kDenseCap || hipMallocAsync(&ws, ws_bytes, stream) != hipSuccess) {
      (void)hipGetLastError();
      ws = nullptr;
      if (ngq_barc > kMaxBarc) return PYGW_E_ARG;
      ws_bytes = minm_is_ws_bytes(ngq, nvi, nvj, ndim);
      HIP_TRY(hipMallocAsync(&ws, ws_bytes, stream));
    }
    own = true;
  }
  int rc = minm_is_launch(nstart, nend, mstart, mend, Aeigk, ngi, Aeigq, ngj, iumklap, mpwipw, ngq, ngq_barc, nvi, nvj,
                          indgk_i, indgk_j, indgq_inverse, inv_size, gindex, ng, i_g0, n1, n2, n3, vol, mmat, ld_mmat,
                          col0, ws, ws_bytes, stream);
  if (own) {
    hipError_t e = hipFreeAsync(ws, stream);
    if (e != hipSuccess && rc == PYGW_OK) rc = PYGW_E_HIP;
  }
  return rc;
}

extern "C" int pygw_pw_basis_workspace(int ngq, int ngq_barc, size_t* bytes) {
  if (ngq <= 0 || ngq > kMaxMb || ngq_barc < ngq || !bytes) return PYGW_E_ARG;
  *bytes = pw_basis_ws_bytes(ngq, ngq_barc);
  return PYGW_OK;
}

extern "C" int pygw_pw_basis(const int* indgq, int ngq, int ngq_barc, const void* ipwint, int ng, const int* gindex,
                             const int* i_g0, int n1, int n2, int n3, void* mpwipw, void* workspace, size_t ws_bytes,
                             void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (!indgq || !ipwint || !gindex || !i_g0 || !mpwipw || ngq <= 0 || ngq > kMaxMb || ngq_barc < ngq || ng <= 0 ||
      n1 < 0 || n2 < 0 || n3 < 0)
    return PYGW_E_ARG;
  void* ws = workspace;
  bool own = false;
  if (!ws) {
    ws_bytes = pw_basis_ws_bytes(ngq, ngq_barc);
    HIP_TRY(hipMallocAsync(&ws, ws_bytes, stream));
    own = true;
  }
  int rc = pw_basis_launch(indgq, ngq, ngq_barc, ipwint, ng, gindex, i_g0, n1, n2, n3, mpwipw, ws, ws_bytes, stream);
  if (own) {
    hipError_t e = hipFreeAsync(ws, stream);
    if (e != hipSuccess && rc == PYGW_OK) rc = PYGW_E_HIP;
  }
  return rc;
}

extern "C" int pygw_coul_pw_block(const void* mpwipw, int ngq, int ngq_barc, const double* gqlen, int q_is_zero,
                                  void* vmat, long ld_vmat, int row0, void* stream) {
  if (!mpwipw || !gqlen || !vmat || ngq <= 0 || ngq_barc <= 0 || row0 < 0 || ld_vmat < (long)row0 + ngq)
    return PYGW_E_ARG;
  return coul_pw_block_launch(mpwipw, ngq, ngq_barc, gqlen, q_is_zero, vmat, ld_vmat, row0, (hipStream_t)stream);
}

extern "C" int pygw_mixed_coulomb_workspace(int max_mult, int ngq_barc, size_t* bytes) {
  if (max_mult <= 0 || ngq_barc <= 0 || !bytes) return PYGW_E_ARG;
  *bytes = mixed_coulomb_ws_bytes(max_mult, ngq_barc);
  return PYGW_OK;
}

extern "C" int pygw_mixed_coulomb(const double* vq, int iat, int q_vq, const double* jlam, int ngs, const int* gindex,
                                  const int* indgq, const double* gqlen0, const int* G_unique, int ngq_barc,
                                  const int* big_l, int nmix, const double* k2cartes, const double* rotloc,
                                  const double* trotij, const int* mult, const double* pos, int nat, double vol,
                                  void* out, long ld_out, int row0, void* workspace, size_t ws_bytes, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (!vq || !jlam || !gindex || !indgq || !gqlen0 || !G_unique || !big_l || !k2cartes || !rotloc || !trotij ||
      !mult || !pos || !out || ngs <= 0 || ngq_barc <= 0 || nmix <= 0 || nat <= 0 || iat < 0 || iat >= nat ||
      !(vol > 0.0) || row0 < 0 || ld_out < ngq_barc)
    return PYGW_E_ARG;
  const size_t need = mixed_coulomb_ws_bytes(mult[iat] > 0 ? mult[iat] : 1, ngq_barc);
  void* ws = workspace;
  bool own = false;
  if (!ws) {
    HIP_TRY(hipMallocAsync(&ws, need, stream));
    own = true;
  } else if (ws_bytes < need) {
    return PYGW_E_ARG;
  }
  int rc = mixed_coulomb_launch(vq, iat, q_vq, jlam, ngs, gindex, indgq, gqlen0, G_unique, ngq_barc, big_l, nmix,
                                k2cartes, rotloc, trotij, mult, pos, nat, vol, out, ld_out, row0, ws, stream);
  if (own) {
    hipError_t e = hipFreeAsync(ws, stream);
    if (e != hipSuccess && rc == PYGW_OK) rc = PYGW_E_HIP;
  }
  return rc;
}

extern "C" int pygw_coul_mixed_block(const void* mpwipw, int ngq, int ngq_barc, const void* vmatit, int loct,
                                     void* vmat, long ld_vmat, void* stream) {
  if (!mpwipw || !vmatit || !vmat || ngq <= 0 || ngq_barc <= 0 || loct <= 0 || ld_vmat < (long)loct + ngq)
    return PYGW_E_ARG;
  return coul_mixed_block_launch(mpwipw, ngq, ngq_barc, vmatit, loct, vmat, ld_vmat, (hipStream_t)stream);
}
