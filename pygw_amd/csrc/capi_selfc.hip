// C-ABI for the Sigma_c part of the path (include/pygw_hip.h).
#include "common.h"

int mwm_f64_launch(const void* M, int nrows, int pb, const void* W, int nom, double wkq, int rpk, void* out,
                   hipStream_t stream);
int mwm_mixed_launch(const void* M, int nrows, int pb, const void* W, int nom, double wkq, int rpk, void* ws,
                     size_t ws_bytes, void* out, hipStream_t stream);
size_t mwm_mixed_ws_bytes(int nrows, int pb, int nom);
int selfc_spline_launch(const void* mwm, int nrows, int nb1, int nom, int nb2, const double* x, const double* xp,
                        int nomp, void* out, hipStream_t stream);
int selfc_conv_launch(const double* enk, int nrows, int nb1, int nb2, const void* mwm, int nom, const double* omega,
                      const void* mwmp, int nomp, const double* omp, const double* womp, void* out,
                      hipStream_t stream);
int padecof_launch(const void* zn, const void* gn, int n, int nstate, void* coef, hipStream_t stream);
int padeg_launch(const void* x, int m, const void* zn, const void* coef, int n, int nstate, void* gx,
                 hipStream_t stream);
int pade_qp_launch(const void* sig, int nstate, int nom, const double* omega, int npts, const double* ene,
                   const double* ene0, const double* sigx, const double* vxc, int iop_es, int iop_gw0, double* out,
                   hipStream_t stream);
int selfc_qsum_launch(const void* parts, int nq, long n, double alpha, double beta, void* out, hipStream_t stream);
int gw0_enk_launch(const double* bands, int nband, const double* core, int ncore, const int* jrk, int nitems, int nb2,
                   double* enk, hipStream_t stream);
int tetra_idos_launch(const double* ebnd, int nband, const int* atet, const double* wtet, int ntet,
                      const double* energies, int ne, double* idos, double* dos_out, hipStream_t stream);

extern "C" int pygw_mwm_workspace(int n_k, int rows_per_k, int pb, int n_om, int precision, size_t* bytes) {
  if (n_k <= 0 || rows_per_k <= 0 || pb <= 0 || n_om <= 0 || !bytes || (precision != 0 && precision != 1))
    return PYGW_E_ARG;
  if ((long)n_k * rows_per_k > (1L << 31) - 1) return PYGW_E_ARG;
  *bytes = precision == 0 ? 0 : mwm_mixed_ws_bytes(n_k * rows_per_k, pb, n_om);
  return PYGW_OK;
}

extern "C" int pygw_mwm(const void* M, int n_k, int rows_per_k, int pb, const void* W, int n_om, double wkq,
                        int precision, void* mwm_out, void* workspace, size_t ws_bytes, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (!M || !W || !mwm_out || n_k <= 0 || rows_per_k <= 0 || pb <= 0 || n_om <= 0) return PYGW_E_ARG;
  if ((long)n_k * rows_per_k > (1L << 31) - 1 || n_om > 65535) return PYGW_E_ARG;
  const int nrows = n_k * rows_per_k;
  if (precision == 0) return mwm_f64_launch(M, nrows, pb, W, n_om, wkq, rows_per_k, mwm_out, stream);
  if (precision != 1) return PYGW_E_ARG;
  void* ws = workspace;
  bool own = false;
  if (!ws) {
    ws_bytes = mwm_mixed_ws_bytes(nrows, pb, n_om);
    HIP_TRY(hipMallocAsync(&ws, ws_bytes, stream));
    own = true;
  }
  // a smaller caller workspace walks the frequencies in smaller slabs
  int rc = mwm_mixed_launch(M, nrows, pb, W, n_om, wkq, rows_per_k, ws, ws_bytes, mwm_out, stream);
  if (own) {
    hipError_t e = hipFreeAsync(ws, stream);
    if (e != hipSuccess && rc == PYGW_OK) rc = PYGW_E_HIP;
  }
  return rc;
}

extern "C" int pygw_selfc_spline(const void* mwm, int n_k, int nb1, int nb2, const double* omega, int n_om,
                                 const double* omega_p, int n_omp, void* mwm_p, void* stream) {
  if (!mwm || !omega || !omega_p || !mwm_p || n_k <= 0 || nb1 <= 0 || nb2 <= 0 || n_omp <= 0) return PYGW_E_ARG;
  return selfc_spline_launch(mwm, n_k * nb1, nb1, n_om, nb2, omega, omega_p, n_omp, mwm_p, (hipStream_t)stream);
}

extern "C" int pygw_selfc_convolve(const double* enk, int n_k, int nb1, int nb2, const void* mwm, const double* omega,
                                   int n_om, const void* mwm_p, const double* omega_p, const double* womeg_p,
                                   int n_omp, void* sc_out, void* stream) {
  if (!enk || !mwm || !omega || !omega_p || !womeg_p || !sc_out || n_k <= 0 || nb1 <= 0 || nb2 <= 0 || n_om <= 0)
    return PYGW_E_ARG;
  if (!mwm_p && n_omp != n_om) return PYGW_E_ARG;
  if ((long)n_k * nb1 > 2147483647L || n_om > 65535) return PYGW_E_ARG;
  return selfc_conv_launch(enk, n_k * nb1, nb1, nb2, mwm, n_om, omega, mwm_p, n_omp, omega_p, womeg_p, sc_out,
                           (hipStream_t)stream);
}

extern "C" int pygw_pade_qp(const void* sigc, int n_state, int n_om, const double* omega, int n_pade,
                            const double* ene, const double* sigx, const double* vxc, double* out, void* stream) {
  if (!sigc || !omega || !ene || !out || n_state <= 0 || n_om <= 0) return PYGW_E_ARG;
  return pade_qp_launch(sigc, n_state, n_om, omega, n_pade, ene, nullptr, sigx, vxc, 0, 0, out, (hipStream_t)stream);
}

extern "C" int pygw_padecof(const void* zn, const void* gn, int n, int n_state, void* coef, void* stream) {
  if (!zn || !gn || !coef || n <= 0 || n > 64 || n_state <= 0) return PYGW_E_ARG;
  return padecof_launch(zn, gn, n, n_state, coef, (hipStream_t)stream);
}

extern "C" int pygw_padeg(const void* x, int m, const void* zn, const void* coef, int n, int n_state, void* gx,
                          void* stream) {
  if (!x || !zn || !coef || !gx || m <= 0 || n <= 0 || n > 64 || n_state <= 0) return PYGW_E_ARG;
  return padeg_launch(x, m, zn, coef, n, n_state, gx, (hipStream_t)stream);
}

extern "C" int pygw_pade_qp_gw0(const void* sigc, int n_state, int n_om, const double* omega, int n_pade,
                                const double* ene, const double* ene0, const double* sigx, const double* vxc,
                                int iop_es, int iop_gw0, double* out, void* stream) {
  if (!sigc || !omega || !ene || !out || n_state <= 0 || n_om <= 0) return PYGW_E_ARG;
  return pade_qp_launch(sigc, n_state, n_om, omega, n_pade, ene, ene0, sigx, vxc, iop_es, iop_gw0, out,
                        (hipStream_t)stream);
}

extern "C" int pygw_selfc_qsum(const void* parts, int n_q, long n, double alpha, double beta, void* out,
                               void* stream) {
  if (!parts || !out || n_q <= 0 || n <= 0) return PYGW_E_ARG;
  return selfc_qsum_launch(parts, n_q, n, alpha, beta, out, (hipStream_t)stream);
}

extern "C" int pygw_gw0_enk(const double* bands, int n_irk, int n_band, const double* core, int n_core,
                            const int* jrk, int n_items, int nb2, double* enk, void* stream) {
  if (!bands || !jrk || !enk || n_irk <= 0 || n_band <= 0 || n_core < 0 || n_items <= 0 || nb2 <= n_core ||
      nb2 - n_core > n_band || (n_core > 0 && !core))
    return PYGW_E_ARG;
  return gw0_enk_launch(bands, n_band, core, n_core, jrk, n_items, nb2, enk, (hipStream_t)stream);
}

extern "C" int pygw_tetra_idos(const double* ebnd, int n_k, int n_band, const int* atet, const double* wtet,
                               int n_tet, const double* energies, int n_e, double* idos, double* dos, void* stream) {
  if (!ebnd || !atet || !wtet || !energies || !idos || !dos || n_k <= 0 || n_band <= 0 || n_tet <= 0 || n_e <= 0 ||
      n_e > 65535)
    return PYGW_E_ARG;
  return tetra_idos_launch(ebnd, n_band, atet, wtet, n_tet, energies, n_e, idos, dos, (hipStream_t)stream);
}
