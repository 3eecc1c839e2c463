/*
 * pygw_hip.h -- C ABI of libpygw_hip.so, the MI355X (gfx950) backend of
 * PyGW's polarization -> dielectric inversion -> Sigma_c path.
 *
 * Conventions for every entry point:
 *   - array arguments are DEVICE pointers (HIP global memory), caller-owned;
 *   - integer index arrays are int32 and 0-based, as PyGW passes them to
 *     its Fortran (for_tetrahedra.f90:445-456 adds the 1);
 *   - `stream` is a hipStream_t (NULL = default stream);
 *   - the return value is 0 on success, < 0 on error:
 *       -10 invalid argument, -11 HIP runtime error,
 *       -12 geometry/classification error (where PyGW's Fortran would `stop`),
 *       -13 matrix not Hermitian positive definite,
 *       -14 eigensolver not converged,
 *       -15 sizes beyond the fixed scratch array of pygw_eps_inverse_noalloc;
 *   - where an `info` pointer is given (host memory), the call synchronises
 *     `stream` and fills it; pass NULL to stay asynchronous;
 *   - `workspace` may be NULL (the library then allocates on `stream`);
 *     otherwise it must hold the number of bytes the *_workspace query
 *     returns.
 * All floating-point outputs are FP64 (complex outputs interleaved re,im).
 */
#ifndef PYGW_HIP_H
#define PYGW_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Library version (major*10000 + minor*100 + patch). */
int pygw_version(void);

/* ------------------------------------------------------------------------
 * Tetrahedron polarization weights.
 * Replaces ft.bz_calcqdepw_par2(enk, ef, kqid, ncg, nomx, numin, omega,
 * tetc) (for_tetrahedra.f90:390-464; called at pygw.py:3443).
 *   enk    [n_bnd, n_k]  Fortran order (band fastest), Hartree
 *   kqid   [n_k]         index of k-q for each k (kqid[:, iq])
 *   tetc   [4, n_tet]    Fortran order, n_tet = 6*n_k
 *   omega  [n_om]        Matsubara frequencies (Hartree)
 *   kcw    layout 0: [n_o, n_e, n_k, n_om] Fortran order (PyGW's array),
 *          layout 1: [n_k][n_om][n_o*n_e] with pair p = i_o*n_e + i_e
 *          (the pair order of pygw.py:3170,3182), where
 *          n_o = ncg+nomx+1, n_e = n_bnd-ncg-numin.
 *   compat bit 0: reproduce PyGW's class-4 vertex rotation
 *                 (for_tetrahedra3.f90:482-486; weights of corners 2 and 4
 *                 exchanged); bit 1: reproduce the wtemp(4,2,nom) aliasing
 *                 of the nnod = 7, 8 polyhedra (for_tetrahedra3.f90:33,
 *                 105-118).  3 = PyGW's results, 0 = the exact integrals.
 *   info   (optional) [2]: status bits (1 geometry error, 2 nnod=5
 *          bipyramid, 4 malformed tetc), number of check_weight warnings.
 * ------------------------------------------------------------------------ */
/* Workspace of pygw_tetra_weights: the k incidence lists plus a buffer of
 * corner weights (n_tet*4*n_om doubles per band pair) for a chunk of pairs;
 * the returned size holds up to 2 GiB of it.  A smaller workspace (at least
 * one chunk of 64 pairs, or of all n_pairs = n_o*n_e when fewer) runs more
 * chunks; NULL workspace: the call allocates the returned size itself. */
int pygw_tetra_weights_workspace(int n_k, int n_pairs, int n_om, size_t* bytes);

/* SVD-basis weights: replaces ft.bz_calcqdepw_par3 (for_tetrahedra.f90:466-542,
 * called at pygw.py:3441) as kcw_l = kcw . dUl applied to the layout-1
 * output of pygw_tetra_weights.  dUl Fortran (n_om, n_l) as built by
 * svd_functions (pygw.py:3478-3501).  out layout 0: Fortran
 * [n_o, n_e, n_k, n_l]; layout 1: [n_k][n_l][n_o*n_e]. */
int pygw_tetra_weights_svd(const double* kcw_native, int n_k, int n_om, int n_o, int n_e, const double* dUl,
                           int n_l, int layout, double* out, void* stream);
int pygw_tetra_weights(const double* enk, double ef, const int* kqid, const int* tetc, int ncg, int nomx,
                       int numin, const double* omega, int n_om, int n_k, int n_bnd, int n_tet, int layout,
                       int compat, double* kcw, void* workspace, size_t ws_bytes, int* info, void* stream);

/* ------------------------------------------------------------------------
 * chi0 contraction (calc_eps, pygw.py:3161-3200):
 *   eps[I,J,w] (+)= sum_k sum_p coef*kcw[k][w][p] * M[k][p][I] * conj(M[k][p][J])
 * Replaces the per-(k, w) `la.matmul(tmat, cminm2)` ZGEMMs with
 * tmat = minm2.T*coef*kcw (pygw.py:3182-3185).
 *   M     [n_k][n_pairs][pb], k stride `stride_k` elements, holding the
 *         k-points k0..k0+n_k-1 (PyGW's minm2 = reshape(minm,
 *         (nb1*nb2, matsiz)) per k, v^1/2 folded in as at pygw.py:2666);
 *         complex128 for precision 0, complex64 for precision 1
 *   kcw   [N_k][n_om][n_pairs] (layout 1 of pygw_tetra_weights), all k
 *   eps   complex128 Fortran [pb, pb, n_om]; beta = 0 overwrite, 1 add
 *   precision 0: FP64 on v_mfma_f64_16x16x4f64
 *   precision 1: FP16 hi+lo split with per-column and per-frequency
 *                power-of-two scaling, 3 products on
 *                v_mfma_f32_32x32x16_f16, FP32 tile sums folded into the
 *                FP64 eps every 64*s chunks of 32 pairs
 *   workspace: pygw_chi0_workspace bytes for (pb, n_pairs, n_om, n_k) (0 for
 *         precision 0): scales, 64 MiB of part slots for the split last
 *         round, and the fragment planes of up to 1 GiB of k-points.  NULL
 *         lets the call allocate it (stream-ordered).  A smaller caller
 *         workspace (at least the scales and one k-point's fragment planes)
 *         makes the call walk k in smaller slabs, and without room for the
 *         part slots it runs the last round unsplit.
 * ------------------------------------------------------------------------ */
int pygw_chi0_workspace(int pb, int n_pairs, int n_om, int n_k, int precision, size_t* bytes);
/* pygw_chi0_launch_timer / pygw_chi0_launch_time: measurement only, no
 *   counterpart in the reference.  Arms HIP events around the next
 *   max_launches launches of the split-precision contraction kernel (0
 *   disarms); pygw_chi0_launch_time synchronises them and returns the summed
 *   kernel time (ms) and the launch count since the last read.  Not for
 *   work under HIP graph capture. */
int pygw_chi0_launch_timer(int max_launches);
int pygw_chi0_launch_time(double* ms, int* launches);
int pygw_chi0_accumulate(const void* M, long stride_k, int n_pairs, int pb, const double* kcw, int n_om, int k0,
                         int n_k, double coef, int precision, int beta, void* eps, void* workspace,
                         size_t ws_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Contraction in a compressed frequency basis.  PyGW's SVD mode (iopfreq
 * 5: svd_functions, pygw.py:3478-3501; bz_calcqdepw_par3,
 * for_tetrahedra.f90:466-542; calc_eps, pygw.py:3206-3226) computes the
 * weights in n_l basis functions, contracts chi0 with n_l "frequencies" and
 * expands eps = dot(eps, Ul) before the inversion.  The same three steps here,
 * with either PyGW's analytic basis (dUl, Ul) or a basis taken from the q's
 * own weights:
 *   pygw_omega_basis: U = eigenvectors of G = sum_{k,p} kcw[k,:,p]
 *     kcw[k,:,p]^T (FP64, cyclic Jacobi) in decreasing order of eigenvalue,
 *     sv[r] = sqrt(lambda_r) (the singular values of the weights as an
 *     (N_k N_p) x n_om matrix); info[0] = Jacobi sweeps.  balance 0:
 *     proj = expand = U; balance 1: G is first scaled to unit diagonal
 *     (every frequency's column of weights to unit norm, D their norms),
 *     U its eigenvectors, proj = D^-1 U, expand = D U (both [n_om][n_om],
 *     element (w, r) at w*n_om + r; proj^T expand = 1), sv of the scaled
 *     matrix.  n_om <= PYGW_OMEGA_MAX.  workspace:
 *     pygw_omega_basis_workspace bytes or NULL (the call allocates it,
 *     stream-ordered).
 *   pygw_omega_project: kcw_r[k][r][p] = sum_w P(w, r) kcw[k][w][p] with
 *     P(w, r) = proj[w*sw + r*sr] (proj above: sw = n_om, sr = 1; Fortran dUl
 *     (n_om, n_l): sw = 1, sr = n_om).  kcw_r then goes to
 *     pygw_chi0_accumulate with n_om = rank.
 *   pygw_omega_expand: eps[w] = beta*eps[w] + sum_r E(w, r) X[r] for the
 *     complex128 [rank][pb][pb] result X of that contraction, E(w, r) =
 *     ex[w*sw + r*sr] (expand above; PyGW's Ul [n_l][n_om]: sw = 1,
 *     sr = n_om); eps complex128 [n_om][pb][pb].
 *   pygw_omega_expand_batched: the same for n_batch blocks of nel complex
 *     values: out[b][w][e] = beta*out[b][w][e] + sum_r E(w, r) x[b][r][e]
 *     (e.g. the MWM of W in a frequency basis, [n_k][rank][rows] ->
 *     [n_k][n_om][rows]; PyGW's SVD mode keeps mwm in the basis and
 *     convolves the basis functions instead, mcommon.py:116-123).
 * The basis of W = eps^-1 - 1 (PyGW: eps = dot(eps, dUl), pygw.py:3273-3276)
 * is pygw_omega_basis of the complex128 [n_om][pb][pb] eps read as real
 * [1][n_om][2 pb^2] (the real part of the Gram matrix of W), and its
 * projection pygw_omega_project of the same real view.
 * ------------------------------------------------------------------------ */
#define PYGW_OMEGA_MAX 96
int pygw_omega_basis_workspace(int n_k, int n_om, int n_pairs, size_t* bytes);
int pygw_omega_basis(const double* kcw, int n_k, int n_om, int n_pairs, int balance, double* proj, double* expand,
                     double* sv, int* info, void* workspace, size_t ws_bytes, void* stream);
/* The same for n_q weight arrays at once (one eigen-decomposition launch of
 * n_q blocks): kcw, n_k, n_pairs are HOST arrays of n_q device pointers /
 * sizes ([n_k[q]][n_om][n_pairs[q]] each); proj, expand [n_q][n_om][n_om],
 * sv [n_q][n_om], info [n_q] (optional) on the device; workspace:
 * pygw_omega_basis_many_workspace bytes or NULL.  Capturable in a HIP graph
 * (the host arrays are read when the call is made). */
int pygw_omega_basis_many_workspace(int n_q, int n_om, size_t* bytes);
int pygw_omega_basis_many(int n_q, const double* const* kcw, const int* n_k, const int* n_pairs, int n_om,
                          int balance, double* proj, double* expand, double* sv, int* info, void* workspace,
                          size_t ws_bytes, void* stream);
int pygw_omega_project(const double* kcw, int n_k, int n_om, int n_pairs, const double* proj, long sw, long sr,
                       int rank, double* kcw_r, void* stream);
int pygw_omega_expand(const void* x, const double* ex, long sw, long sr, int pb, int rank, int n_om, int beta,
                      void* eps, void* stream);
int pygw_omega_expand_batched(const void* x, const double* ex, long sw, long sr, int n_batch, long nel, int rank,
                              int n_om, int beta, void* out, void* stream);

/* Autotuner (no counterpart in PyGW): times the contraction's schedules
 * s = 1, 2, 4 (FP64: s Matsubara frequencies per 64x64 tile block; split
 * precision: FP32 tile sums folded into FP64 every 64*s chunks) on the given
 * inputs and records the fastest for (device, precision, pb, n_pairs, n_om);
 * later pygw_chi0_accumulate calls with that shape use it.  eps is
 * overwritten (scratch).  best: chosen frequencies per block; times[3]
 * (optional): milliseconds of each candidate (-1: not tried; split
 * precision tries s = 1, 2 only, see the batch API below). */
int pygw_chi0_autotune(const void* M, long stride_k, int n_pairs, int pb, const double* kcw, int n_om, int k0,
                       int n_k, double coef, int precision, void* eps, int* best, float* times, void* stream);
/* Sets (nw = 1, 2, 4) or clears (nw = 0) the cached schedule of one shape. */
int pygw_chi0_set_schedule(int precision, int pb, int n_pairs, int n_om, int nw);

/* ------------------------------------------------------------------------
 * Batched split-precision chi0 over many q-points (no single counterpart in
 * PyGW: it is the q loop of G0W0.Compute, pygw.py:3654-3718, around the
 * contraction of calc_eps, pygw.py:3161-3200, for uneven PB per q and
 * ragged per-k pair windows).  Per q: M[q] complex64 [K_q][PB_q] (the
 * window pairs of all k concatenated), kcw[q] FP64 [n_om][K_q], eps[q]
 * complex128 Fortran [PB_q, PB_q, n_om] (overwritten).  plan uploads the
 * descriptor tables into the workspace (synchronous on `stream`) and returns
 * an opaque plan; run only launches kernels (capturable in a HIP graph):
 * per-q scales and fragment planes, one contraction launch over all
 * (q, I tile, merged column-frequency chunk) work items, per-q mirror.  schedule 0 = the
 * autotuned one (default 2); 1, 2, 4 = fold every 64*s chunks.  The
 * autotuners (split precision) choose among s = 1, 2 (times[2] = -1): s = 4
 * doubles the FP32 partial sums and lands near the 1e-6 accuracy bar.
 * ------------------------------------------------------------------------ */
int pygw_chi0_batch_workspace(int n_q, const int* pb, const int* n_pairs, int n_om, size_t* bytes);
int pygw_chi0_batch_plan(int n_q, const void* const* M, const int* pb, const int* n_pairs, const double* const* kcw,
                         int n_om, double coef, void* const* eps, void* workspace, size_t ws_bytes, void* stream,
                         void** plan);
int pygw_chi0_batch_run(void* plan, int schedule, void* stream);
int pygw_chi0_batch_autotune(void* plan, int* best, float* times, void* stream);
int pygw_chi0_batch_free(void* plan);

/* ------------------------------------------------------------------------
 * Dielectric inversion (calc_eps, pygw.py:3229-3271, q != 0):
 *   eps[:,:,w] <- ((eps + eps^H)/2 + 1)^{-1} - 1   for every w, in place.
 * Replaces `linalg.inv` (pygw.py:3244); block Gauss-Jordan in FP64 without
 * pivoting (the matrix is Hermitian positive definite).
 *   info (optional) [1]: non-zero if some pivot was not positive
 *   (return -13).
 * ------------------------------------------------------------------------ */
int pygw_eps_inverse_workspace(int pb, int n_om, size_t* bytes);
int pygw_eps_inverse(void* eps, int pb, int n_om, void* workspace, size_t ws_bytes, int* info, void* stream);

/* pygw_eps_inverse_noalloc: pygw_eps_inverse with no allocation at all.  No
 * counterpart in PyGW: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * allocation).  The kernels of pygw_eps_inverse run on a fixed device array
 * of PYGW_NOALLOC_SCRATCH_BYTES that the library holds on every device it
 * runs on until the process ends; the call needs pygw_eps_inverse_workspace
 * (pb, n_om) <= that size (about PB 900 with 68 frequencies, PB 1200 with 51)
 * and returns -15 otherwise, before any launch.  Calls on one device share
 * its array: calls that overlap on one device (two streams or threads) race.
 * info as for pygw_eps_inverse (NULL keeps the call asynchronous). */
#define PYGW_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int pygw_eps_inverse_noalloc(void* eps, int pb, int n_om, int* info, void* stream);

/* ------------------------------------------------------------------------
 * MWM contraction (calc_selfc, pygw.py:3353-3358):
 *   mwm[k][w][r] = wkq * sum_{I,J} conj(M[k][r][I]) W[I,J,w] M[k][r][J]
 * Replaces `la.matmul(cminm2, eps[:,:,iom])` and the row sum.
 *   M    [n_k][rows_per_k][pb], rows r = ib1*nb2 + ib2 (minm2 of the
 *        self-energy bands, pygw.py:3349-3351); complex128 for precision 0,
 *        complex64 for precision 1
 *   W    complex128 Fortran [pb, pb, n_om] (eps after pygw_eps_inverse)
 *   out  complex128 [n_k][n_om][rows_per_k] (PyGW's mwm[iom, ib1, ib2] per k)
 *   precision 0 = FP64 MFMA; 1 = FP16 hi+lo split (rows of M and columns
 *   of W scaled by powers of two), 3 products, FP32 accumulation over I,
 *   FP64 row sum over J
 *   workspace: pygw_mwm_workspace bytes (0 for precision 0); NULL lets the
 *   call allocate it (stream-ordered); a smaller one (at least the row
 *   planes and one frequency's W planes) walks the frequencies in slabs.
 * ------------------------------------------------------------------------ */
int pygw_mwm_workspace(int n_k, int rows_per_k, int pb, int n_om, int precision, size_t* bytes);
int pygw_mwm(const void* M, int n_k, int rows_per_k, int pb, const void* W, int n_om, double wkq, int precision,
             void* mwm_out, void* workspace, size_t ws_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Spline refinement of mwm onto the precise mesh (Compute_selfc_inside,
 * mcommon.py:126-140): cubic spline of (w^2+1) mwm with y''(w_0) = 0 and
 * y'(w_max) = 0, extrapolated at the ends, divided by (w_p^2+1).
 *   mwm    complex128 [n_k][n_om][nb1][nb2]
 *   mwm_p  complex128 [n_k][nb1][nb2][n_omp]
 * ------------------------------------------------------------------------ */
int pygw_selfc_spline(const void* mwm, int n_k, int nb1, int nb2, const double* omega, int n_om,
                      const double* omega_p, int n_omp, void* mwm_p, void* stream);

/* ------------------------------------------------------------------------
 * G0 (x) W convolution on the imaginary axis.  Replaces
 * fnc.all_fr_convolutions (mwm_p = NULL, coarse mesh; fnc.f90:139-159) and
 * the loop over fnc.few_fr_convolutions (mcommon.py:141-144, fnc.f90:161-176).
 *   enk    [n_k][nb2] energies of the k-q states (Hartree)
 *   sc_out complex128 [n_k][nb1][n_om] (PyGW's sc_p[irk, ie1, iom])
 * ------------------------------------------------------------------------ */
int pygw_selfc_convolve(const double* enk, int n_k, int nb1, int nb2, const void* mwm, const double* omega,
                        int n_om, const void* mwm_p, const double* omega_p, const double* womeg_p, int n_omp,
                        void* sc_out, void* stream);

/* ------------------------------------------------------------------------
 * Pade continuation and quasiparticle energies (AnalyticContinuation
 * iop_ac = 0, mcommon.py:330-338, with fpade.padecof/padeg,
 * for_pade.f90:3-51; Compute_quasiparticles G0W0 update, mcommon.py:182-203).
 *   sigc  complex128 [n_state][n_om]; the first n_pade points are used
 *   ene, sigx, vxc  [n_state] (sigx, vxc may be NULL = 0)
 *   out   [n_state][6]: Re S(e), Im S(e), Re dS/de, Im dS/de, Z, E_qp
 * ------------------------------------------------------------------------ */
int pygw_pade_qp(const void* sigc, int n_state, int n_om, const double* omega, int n_pade, const double* ene,
                 const double* sigx, const double* vxc, double* out, void* stream);

/* ------------------------------------------------------------------------
 * Pade continuation through arbitrary complex points.
 * pygw_padecof replaces fpade.padecof(gn, zn) (for_pade.f90:3-25) for
 *   n_state functions on common nodes: gn complex128 [n_state][n], zn
 *   complex128 [n] (n <= 64), coef complex128 [n_state][n] (Thiele's
 *   reciprocal differences p(j, j)).
 * pygw_padeg replaces fpade.padeg(x, zn, coef) (for_pade.f90:30-51): the
 *   continued fraction of each function at x complex128 [m]; gx complex128
 *   [n_state][m].  (mcommon.py:336-338 calls the pair with zn = i omega_n.)
 * ------------------------------------------------------------------------ */
int pygw_padecof(const void* zn, const void* gn, int n, int n_state, void* coef, void* stream);
int pygw_padeg(const void* x, int m, const void* zn, const void* coef, int n, int n_state, void* gx, void* stream);

/* ------------------------------------------------------------------------
 * GW0 on stored mwm (the calceqp iteration, postprocess.py:206-321: Sigma_c
 * of every (q, irk) recomputed with the quasiparticle bands, mixed, then the
 * quasiparticle energies and the Fermi level again).
 *
 * pygw_pade_qp_gw0 replaces Compute_quasiparticles (mcommon.py:167-210) for
 *   both of its modes: iop_es = 0 is pygw_pade_qp (G0W0, ene0 unused);
 *   iop_es = -1 is the GW0 update with ene0 the Kohn-Sham energies and
 *   ene the current bands: iop_gw0 = 1: E = e0 + s - vxc; 2: E = e + Z(s -
 *   vxc) + e0 - e; otherwise E = e + Z(s - vxc + e0 - e) (s = Re S(e) +
 *   sigx).  out as pygw_pade_qp.
 * pygw_selfc_qsum: out = beta*out + alpha*sum_q parts[q] over n_q blocks of
 *   n complex128 values, summed in q order (the q loop of Compute_selfc,
 *   postprocess.py:176-186, and the mixing sigc*(1-mix) + mix*sigc_new,
 *   postprocess.py:297).  beta = 0 does not read out.
 * pygw_gw0_enk: enk [n_items][nb2] of Compute_selfc_inside (mcommon.py:
 *   94-101): core energies, then bands[jrk[item]][0 .. nb2 - n_core - 1];
 *   bands [n_irk][n_band] row-major, jrk the irreducible index of k - q of
 *   each (q, irk) item (0-based).
 * pygw_tetra_idos replaces ft.idos / ft.dostet (for_tetrahedra.f90:84-141)
 *   at n_e energies at once: idos[e] = sum wtet*intdos1t, dos[e] = sum
 *   wtet*dos1t over bands and tetrahedra; ebnd [n_k][n_band] row-major,
 *   atet [n_tet][4] (0-based), wtet [n_tet].
 * ------------------------------------------------------------------------ */
int pygw_pade_qp_gw0(const void* sigc, int n_state, int n_om, const double* omega, int n_pade, const double* ene,
                     const double* ene0, const double* sigx, const double* vxc, int iop_es, int iop_gw0, double* out,
                     void* stream);
int pygw_selfc_qsum(const void* parts, int n_q, long n, double alpha, double beta, void* out, void* stream);
int pygw_gw0_enk(const double* bands, int n_irk, int n_band, const double* core, int n_core, const int* jrk,
                 int n_items, int nb2, double* enk, void* stream);
int pygw_tetra_idos(const double* ebnd, int n_k, int n_band, const int* atet, const double* wtet, int n_tet,
                    const double* energies, int n_e, double* idos, double* dos, void* stream);

/* ------------------------------------------------------------------------
 * Synthetic matrix elements for benchmarks and tests.  No counterpart in
 * PyGW (its M comes from calc_minm, pygw.py:2547-2721); for callers that
 * need seeded inputs of the benchmark shapes without a DFT calculation.
 *   out[k-k0][p][I], k = k0..k0+n_kc-1, p = i_o*n_e + i_e:
 *   complex Gaussian(seed, counter) * exp(-|e_n(k)-e_m(k-q)|/20 eV)
 *   / sqrt(pb) * vhalf[I], n = row_band0 + i_o, m = col_band0 + i_e;
 *   complex128 if complex128 != 0 else complex64.
 * ------------------------------------------------------------------------ */
int pygw_synth_minm(const double* enk, int n_bnd, const int* kqid, int row_band0, int col_band0, int n_o, int n_e,
                    int pb, const double* vhalf, unsigned long long seed, int k0, int n_kc, int complex128,
                    void* out, void* stream);
/* The same draws for a list of (k, p) pairs (int32 pairs[2*j] = k, pairs[2*j+1] = p):
 * out[j][I] = M[k_j][p_j][I]; for ragged per-k pair windows. */
int pygw_synth_minm_pairs(const double* enk, int n_bnd, const int* kqid, int row_band0, int col_band0, int n_o,
                          int n_e, int pb, const double* vhalf, unsigned long long seed, const int* pairs,
                          long n_pairs, int complex128, void* out, void* stream);

/* ------------------------------------------------------------------------
 * q = 0 (Gamma point) head and wings.
 *
 * Wings: the iq == 0 branch of calc_eps, wtmp = dot(tmat, pm) added to
 * epsw1[:, iom] (pygw.py:3177-3193):
 *   epsw1[I, w] (+)= sum_k sum_p coef*kcw[k][w][p] * M[k][p][I] * pm[k][p]
 *   M as in pygw_chi0_accumulate (complex64 = 1 for complex64 M, else
 *   complex128), kcw kernel-native [N_k][n_om][n_pairs] read from k-point k0,
 *   pm complex128 [n_k][n_pairs], epsw1 complex128 Fortran [pb, n_om]
 *   (beta = 0 overwrites, 1 accumulates).  NULL workspace: allocated.
 * ------------------------------------------------------------------------ */
int pygw_chi0_wings_workspace(int pb, int n_om, int n_k, size_t* bytes);
int pygw_chi0_wings(const void* M, long stride_k, int n_pairs, int pb, int complex64, const double* kcw, int n_om,
                    int k0, int n_k, double coef, const void* pm, void* epsw1, int beta, void* workspace,
                    size_t ws_bytes, void* stream);

/* Inversion with the head/wings update of iq == 0 (pygw.py:3202-3263):
 * as pygw_eps_inverse, and between the inverse A and the final -1:
 * epsw2 = conj(epsw1); bw1 = A epsw1; w2b = conj(A epsw2);
 * ww = dot(epsw2, bw1); emac[w] = (1 + head - ww, 1 + head);
 * h = 1/(1 + head - ww); epsw1 = -h bw1; epsw2 = -h w2b;
 * A += h bw1 w2b^T; head = h - 1.
 *   head complex128 [n_om] (in/out), epsw1 Fortran [pb, n_om] (in/out),
 *   epsw2 (out, same shape), emac complex128 [n_om][2] (out). */
int pygw_eps_inverse_head_workspace(int pb, int n_om, size_t* bytes);
int pygw_eps_inverse_head(void* eps, int pb, int n_om, void* head, void* epsw1, void* epsw2, void* emac,
                          void* workspace, size_t ws_bytes, int* info, void* stream);

/* q = 0 terms of calc_selfc (pygw.py:3359-3364), added to the MWM output of
 * pygw_mwm (mwm [n_k][n_om][nb1][nb2]) at (b, b + ncg_c) for b < nb1:
 *   coefs2*head[w] + coefs1*sum_I M[b, b+ncg_c, I] epsw2[I, w]
 *                  + coefs1*sum_I conj(M[b, b+ncg_c, I]) epsw1[I, w]
 * M_selfe as in pygw_mwm ([n_k][nb1*nb2][pb], complex64 or complex128);
 * coefs1 = singc1*sqrt(4 pi/Vol), coefs2 = singc2*4 pi/Vol (pygw.py:3310-3311). */
int pygw_mwm_head(const void* M_selfe, int n_k, int nb1, int nb2, int pb, int ncg_c, int complex64, const void* head,
                  const void* epsw1, const void* epsw2, int n_om, double coefs1, double coefs2, void* mwm,
                  void* stream);

/* ------------------------------------------------------------------------
 * The v^(1/2) basis of the Coulomb matrix and the matrix elements in it.
 *
 * pygw_coul_setev replaces linalg.eigh(Vmat) (pygw.py:2440), the q = 0
 *   choice of immax (pygw.py:2447-2454) and Coul_setev with
 *   Real_sqrt_Vcoul = False (pygw.py:18, 2505-2545).
 *   vmat: complex128 C order [mbsize][mbsize], Hermitian; only its lower
 *   triangle is read (like eigh).  Eigenvalues and eigenvectors come from a
 *   block cyclic Jacobi method on the device.
 *   - ev [mbsize]: ascending (the eigh order).
 *   - vecs: optional, NULL or Fortran [mbsize, mbsize] eigenvectors in that
 *     order (PyGW's self.Vmat).
 *   - wi0 (q = 0 only, else NULL): complex128 [mbsize] overlaps of the basis
 *     with the constant function.  *immax is the eigenvector (eigh order)
 *     maximising |sum_i wi0[i] conj(V[i][j])|; it is dropped
 *     (iop_coul_x = 0) or kept whatever its eigenvalue (iop_coul_x != 0).
 *   - The other eigenvectors are kept when ev > evtol.
 *   - *matsize = the kept count (PB).
 *   - barcev [mbsize] (first matsize used): kept eigenvalues, decreasing.
 *   - barcvm: Fortran [mbsize, matsize], leading dimension mbsize, with
 *     mbsize columns allocated; column m = v_j sqrt|ev_j|.
 *   Eigenvector phases (and bases of degenerate eigenspaces) differ from
 *   LAPACK's; the products downstream do not depend on them.
 *   The call synchronises the stream once per sweep.
 *   - *n_sweeps (optional): the sweeps run.
 *   - Returns PYGW_E_NOCONV after 60 sweeps without convergence.
 *   workspace: NULL (allocated) or pygw_coul_setev_workspace bytes.
 *   mbsize <= 16384.
 * pygw_minm_vhalf replaces minm = la.matmul(mmat, conj(barcvm))
 *   (pygw.py:2666) and minc = la.matmul(mmat_c, conj(barcvm[:loctmatsize,
 *   :])) (pygw.py:2684):
 *     minm[r][m] = sum_{i < n_in} mmat[r*ld_mmat + i] conj(barcvm[m*ld_barcvm + i]).
 *   mmat: complex128, rows of ld_mmat >= n_in.
 *   minm: row-major [n_rows][matsize], complex128 or (complex64_out = 1)
 *   the FP64 result rounded to complex64 (the operand of the
 *   split-precision chi0).  FP64 MFMA.
 * ------------------------------------------------------------------------ */
/* pygw_pw_basis replaces OrthogonalizedBasisInterstitials (pygw.py:
 *   2126-2214, with PyGW's defaults Real_sqrt_olap = True and FORT = True,
 *   i.e. cmp_pw_olap2, for_Coulomb.f90:306-325):
 *     mpwipw = olap^(-1/2) tmat, olap[j][i] = ipwint[G_i - G_j] (i, j < ngq),
 *     tmat[j][i] = ipwint[G_i - G_j] (j < ngq, i < ngq_barc),
 *   G_i = gindex[indgq[i]], the lookup through i_g0 ([2n1+1][2n2+1][2n3+1],
 *   C order, -1 = none); olap^(-1/2) from the device Jacobi eigensolver.
 *   indgq [ngq_barc], gindex [ng][3], i_g0 int32 0-based; ipwint complex128
 *   [ng]; mpwipw complex128 [ngq][ngq_barc] out.  A difference outside the
 *   i_g0 box gives 0 (cmp_pw_olap2 reuses the previous pair's index there).
 *   Returns PYGW_E_NOTPD if olap has a zero eigenvalue.  Synchronises.
 * pygw_coul_pw_block replaces the plane-wave block of Coulomb (pygw.py:
 *   2419-2430): vmat[(row0 + i) ld + row0 + k] = sum_j mpwipw[i][j] 4 pi /
 *   gqlen[j]^2 conj(mpwipw[k][j]), gqlen float64 [ngq_barc] in device
 *   memory (|q + G| of the basis), gqlen[0] taken as 1e100 when q_is_zero;
 *   row0 = loctmatsize. */
int pygw_pw_basis_workspace(int ngq, int ngq_barc, size_t* bytes);
int pygw_pw_basis(const int* indgq, int ngq, int ngq_barc, const void* ipwint, int ng, const int* gindex,
                  const int* i_g0, int n1, int n2, int n3, void* mpwipw, void* workspace, size_t ws_bytes,
                  void* stream);
int pygw_coul_pw_block(const void* mpwipw, int ngq, int ngq_barc, const double* gqlen, int q_is_zero, void* vmat,
                       long ld_vmat, int row0, void* stream);
/* pygw_mixed_coulomb replaces fCl.mixed_coulomb(vq, iat+1, True, jlam, ...)
 *   (pygw.py:2396, for_Coulomb.f90:327-408) for the equivalent atoms of atom
 *   iat (0-based): rows row0 + im of out (ld_out >= ngq_barc), im over
 *   (equivalent atom, product function irm, M) in the reference's order:
 *     Y_LM(R (q + G_ig)) e^{-2 pi i pos.G_ig} (4 pi/sqrt(vol)) (-i)^L
 *     jlam[irm][ipw] (4 pi/gqlen0[ipw]^2 if q_vq),
 *   ipw the group of ig (G_unique: 0-based starts of the equal-|q+G| groups
 *   of the basis), the G = 0 column 0 at q = 0.  Device: jlam float64
 *   [nmix][ngs], gindex int32 [ng][3], indgq int32, gqlen0 float64 [ngs],
 *   G_unique int32 [ngs], out complex128.  Host: vq [3] (lattice units),
 *   big_l [nmix] (L <= 8, nmix <= 64), k2cartes [3][3], rotloc [nat][3][3],
 *   trotij [ndf][3][3], mult [nat] (<= 24 per atom), pos [ndf][3]
 *   (fractional); R = rotloc[iat] trotij[idf]^T k2cartes.
 * pygw_coul_mixed_block replaces mat2 = dot(mpwipw, Vmatit.T) and its two
 *   placements (pygw.py:2409-2416): vmat[(loct + i) ld + im] = mat2[i][im],
 *   vmat[im ld + loct + i] = conj(mat2[i][im]); vmatit complex128
 *   [loct][ngq_barc], vmat [loct + ngq][ld].
 * pygw_mixed_coulomb_workspace: bytes of the Y_lm scratch for atoms of up to
 *   max_mult equivalent atoms (max_mult * ngq_barc * 81 complex128); the
 *   workspace of pygw_mixed_coulomb must hold it for mult[iat] (NULL: the
 *   call allocates it on `stream`). */
int pygw_mixed_coulomb_workspace(int max_mult, int ngq_barc, size_t* bytes);
int pygw_mixed_coulomb(const double* vq, int iat, int q_vq, const double* jlam, int ngs, const int* gindex,
                       const int* indgq, const double* gqlen0, const int* G_unique, int ngq_barc, const int* big_l,
                       int nmix, const double* k2cartes, const double* rotloc, const double* trotij, const int* mult,
                       const double* pos, int nat, double vol, void* out, long ld_out, int row0, void* workspace,
                       size_t ws_bytes, void* stream);
int pygw_coul_mixed_block(const void* mpwipw, int ngq, int ngq_barc, const void* vmatit, int loct, void* vmat,
                          long ld_vmat, void* stream);
/* The muffin-tin block of the Coulomb matrix and the q = 0 term
 * (mt_coulomb.hip).  All arrays on the device, Fortran layouts as PyGW
 * passes them:
 * pygw_ewald_summation replaces fCl.ewald_summation(cq, lam_max, vpos, br2,
 *   rbas, alat, Vol, ortho, r_cf, g_cf, eta) (pygw.py:2350,
 *   for_Coulomb.f90:1-164).  The lattice arguments are in HOST memory, as
 *   PyGW passes them: pq [3] cartesian q, pos [3][ndf] (Fortran pos(3, ndf),
 *   fractional), br2 (3,3) (columns: reciprocal vectors), rbas (3,3) (rows:
 *   lattice vectors), alat [3]; ortho 0/1; r_cf and g_cf below 64 shortest
 *   basis vectors, lam_max <= PYGW_EWALD_MAX_LAM, ndf <= 1000.  The host lists the lattice
 *   vectors as genr_str does and uploads them into the workspace
 *   (pygw_ewald_workspace bytes for the same lattice arguments, or NULL);
 *   the call synchronises the stream once before it returns.  sgm (device)
 *   complex128 Fortran ((lam_max+1)^2, ndf(ndf+1)/2), column i + j(j+1)/2
 *   for atoms i <= j (0-based).
 * pygw_mt_coulomb replaces fCl.mt_coulomb(vq, iat+1, loctmatsize, big_l,
 *   nmix, im_start, tilg, rrint, djmm, rtlij, sgm, mult, lam_max)
 *   (pygw.py:2394, for_Coulomb.f90:166-283): iat 0-based (the rows of that
 *   atom type's functions, every other row 0) or -1 (all rows: the sum over
 *   iat that PyGW forms); big_l int32 (max_nmix, nat), nmix [nat], im_start
 *   [nat+1] (0-based offsets; also im_start_host in host memory), mult
 *   [nat], tilg [tsize] (calc_tildeg's table), rrint (nat, nijrm), djmm
 *   complex128 (dimdj, ndf), rtlij (max_nmix, max_nmix, nat, nat), sgm from
 *   pygw_ewald_summation with lam_max >= 2 max L; vq [3] (q = 0 leaves the
 *   L = 0 x L = 0 entries 0).  vmat complex128 Fortran (loctmatsize,
 *   loctmatsize), overwritten; an entry whose lookup falls outside tilg,
 *   djmm or sgm is NaN.  Allocates nothing.
 * pygw_radial_olap: out[f][k] = RINT13g(rel, ul_f, us_f, F, F, dh, npt,
 *   rx[0]) (radials.f90:140-187) with F = j_L(k rx) rx, L = lf[f] <= lmax
 *   (mode 0: jlam of Calc_Olap_PW_Product, pygw.py:2110-2124) or
 *   F = sin(k rx) (mode 1, lmax 0, every lf[f] 0: Zero_Momentum_Coulomb's
 *   sing, pygw.py:2069-2086); rx [npt <= 1024] the logarithmic mesh
 *   (strc.radial_mesh, rx[0] e^{j dh}), ul, us float64 [n_fun][npt] (r u and
 *   r us, PyGW's ul_product / us_product), lf int32 [n_fun], k [n_k], out
 *   [n_fun][n_k].  workspace: pygw_radial_olap_workspace bytes (the table
 *   of F and the weights) or NULL.
 * pygw_zero_momentum_coulomb replaces the assembly of Zero_Momentum_Coulomb
 *   (pygw.py:2055-2104) for the L = 0 product functions: the plane waves
 *   gindex int32 (n_g, 3) (Fortran) sorted by length, g_start [n_short+1]
 *   the first index of each run of equal length (pw_get_short_length; the
 *   first plane wave, G = 0, excluded), glen0 [n_short] the first length of
 *   each run, pos (3, ndf) fractional, sing float64 [*][n_short]
 *   [n_sing][n_short] (pygw_radial_olap mode 1 of the L = 0 functions),
 *   and for each L = 0 function a its row f_sing[a] of sing, atom
 *   f_atom[a] and Vmat index f_row[a]:
 *     vmat[f_row[a] + mbsize f_row[b]] = sum_g phase(i, j, g) 16 pi^2/(Vol glen0_g^4)
 *                                        sing[f_sing[a]][g] sing[f_sing[b]][g],
 *   phase(i, j, g) = the sum of e^{-2 pi i G.(pos_i - pos_j)} over the run
 *   (pw_get_short_and_phase, for_Coulomb.f90:515-582).  Only those entries
 *   are written.  workspace: pygw_zero_momentum_workspace bytes or NULL. */
#define PYGW_EWALD_MAX_LAM 16
int pygw_ewald_workspace(const double* pq, int lam_max, const double* pos, const double* br2, const double* rbas,
                         const double* alat, int ortho, double r_cf, double g_cf, int ndf, size_t* bytes);
int pygw_ewald_summation(const double* pq, int lam_max, const double* pos, const double* br2, const double* rbas,
                         const double* alat, double vol, int ortho, double r_cf, double g_cf, double eta, int ndf,
                         void* sgm, void* workspace, size_t ws_bytes, void* stream);
int pygw_mt_coulomb(const double* vq, int iat, int loctmatsize, const int* big_l, const int* nmix,
                    const int* im_start, const int* im_start_host, const int* mult, int nat, int max_nmix,
                    const double* tilg, int tsize, const double* rrint, const void* djmm, int dimdj,
                    const double* rtlij, const void* sgm, int lam_max, void* vmat, void* stream);
int pygw_radial_olap_workspace(int mode, int lmax, int npt, int n_k, size_t* bytes);
int pygw_radial_olap(int mode, int rel, const double* ul, const double* us, const int* lf, int n_fun, int lmax,
                     int npt, const double* rx, double dh, const double* k, int n_k, double* out, void* workspace,
                     size_t ws_bytes, void* stream);
int pygw_zero_momentum_workspace(int ndf, int n_short, size_t* bytes);
int pygw_zero_momentum_coulomb(const int* gindex, int n_g, const int* g_start, const double* glen0, int n_short,
                               const double* pos, int ndf, const double* sing, int n_sing, const int* f_sing,
                               const int* f_atom, const int* f_row, int n_fun0, double vol, int mbsize, void* vmat,
                               void* workspace, size_t ws_bytes, void* stream);
/* pygw_minm_is replaces fvxcn.calc_minm_is (pygw.py:2661,
 *   for_vxcnn.f90:1030-1134), the interstitial part of mmat:
 *     mmat[ie1][ie2][col0 + G] = 1/sqrt(vol) sum_{i < nvi, j < nvj}
 *       Aeigq[mstart+ie2][j] mpwipw[G][K(i, j)] Aeigk[nstart+ie1][i],
 *   K(i, j) = indgq_inverse[i_g0[G_i - G_j + iumklap]] with G_i =
 *   gindex[indgk_i[i]], G_j = gindex[indgk_j[j]]; no term where the
 *   difference leaves the i_g0 box [2n1+1][2n2+1][2n3+1] (C order) or an
 *   index is -1 or out of range.  All index arrays int32, 0-based, on the
 *   device; iumklap: 3 ints in host memory.  Aeigk [nbmax_i][ngi] and Aeigq
 *   [nbmax_j][ngj] complex128 row-major, mpwipw [ngq][ngq_barc] complex128
 *   (ngq_barc <= 4096 in the gathered form; any in the dense form).  mmat complex128, rows (ie1*(mend-mstart) + ie2) of
 *   ld_mmat >= col0 + ngq (col0 = loctmatsize writes into the concatenated
 *   mmat).  FP64 MFMA.  Three associations of the same sum, chosen by
 *   FLOP count and workspace: the pair density rho[(ie1, ie2)][K] first,
 *   then mmat = rho mpwipw^T (8 ndim ngq_barc mdim (nvj + ngq) FLOP), or
 *   T[G][j][ie1] first, then Aeigq T (8 ngq ndim nvj (ngq_barc + mdim)),
 *   the latter as a dense GEMM over the product basis or gathered.
 *   workspace: NULL or pygw_minm_is_workspace bytes (the preferred form
 *   of these sizes); a workspace that holds the dense form's buffers runs
 *   it; one of at least the kmap and T parts (nvi*nvj int32 rounded to
 *   256 B, then ngq*nvj*ndim complex128) runs the gathered form.  The
 *   pair-density form adds its slices over j and over K in a fixed order:
 *   results do not depend on the launch. */
int pygw_minm_is_workspace(int ngq, int ngq_barc, int nvi, int nvj, int ndim, int mdim, size_t* bytes);
int pygw_minm_is(int nstart, int nend, int mstart, int mend, const void* Aeigk, int nbmax_i, int ngi,
                 const void* Aeigq, int nbmax_j, int ngj, const int* iumklap, const void* mpwipw, int ngq,
                 int ngq_barc, int nvi, int nvj, const int* indgk_i, const int* indgk_j, const int* indgq_inverse,
                 int inv_size, const int* gindex, int ng, const int* i_g0, int n1, int n2, int n3, double vol,
                 void* mmat, long ld_mmat, int col0, void* workspace, size_t ws_bytes, void* stream);
int pygw_coul_setev_workspace(int mbsize, size_t* bytes);
int pygw_coul_setev(const void* vmat, int mbsize, double evtol, const void* wi0, int iop_coul_x, double* ev,
                    double* barcev, void* barcvm, void* vecs, int* matsize, int* immax, int* n_sweeps,
                    void* workspace, size_t ws_bytes, void* stream);
int pygw_minm_vhalf(const void* mmat, long n_rows, int n_in, long ld_mmat, const void* barcvm, int ld_barcvm,
                    int matsize, int complex64_out, void* minm, void* stream);
/* pygw_minm_vhalf_mixed: the same product as pygw_minm_vhalf (pygw.py:2666,
 *   2684) in split precision, for the split-precision polarization (which
 *   contracts minm in complex64): mmat's rows and barcvm's columns split
 *   into FP16 hi/lo parts (scaled per row / per column by powers of two),
 *   three MFMA terms per real product, FP32 accumulation over i, the result
 *   unscaled exactly; relative error ~1e-7 of each row's size.
 *   pygw_minm_vhalf_split_basis writes barcvm's split planes and column
 *   scales (once per basis) into `basis` of pygw_minm_vhalf_split_basis_bytes
 *   bytes; pygw_minm_vhalf_mixed then takes them for any n_in <= mbsize,
 *   with a workspace of pygw_minm_vhalf_mixed_workspace bytes (the rows'
 *   planes and scales).  n_rows <= 2^30.  minm as in pygw_minm_vhalf. */
int pygw_minm_vhalf_split_basis_bytes(int mbsize, int matsize, size_t* bytes);
int pygw_minm_vhalf_split_basis(const void* barcvm, int mbsize, int ld_barcvm, int matsize, void* basis,
                                size_t bytes, void* stream);
int pygw_minm_vhalf_mixed_workspace(long n_rows, int n_in, size_t* bytes);
int pygw_minm_vhalf_mixed(const void* mmat, long n_rows, int n_in, long ld_mmat, const void* basis, int mbsize,
                          int matsize, int complex64_out, void* minm, void* workspace, size_t ws_bytes,
                          void* stream);

/* ------------------------------------------------------------------------
 * The muffin-tin and core parts of the matrix elements (minm_mt.hip).
 * pygw_minm_mt_plan (host only) lays out the atom / product-function /
 *   radial-channel structure that PyGW passes to fvxcn.calc_minm_mt as
 *   mult [nat], nmix [nat], big_l (max_nmix, nat), nLO_at (lomax+1, nat),
 *   ncore [nat] (host arrays, Fortran layouts) and in1.lmax: with plan =
 *   NULL it returns the size in *plan_ints, else fills plan (int32).  The
 *   same plan is passed in host memory (plan_host) and device memory
 *   (plan_dev) to the calls below; it does not depend on k or q.
 * pygw_minm_mt replaces fvxcn.calc_minm_mt(ql, nst, nend, mst, mend, alfa,
 *   beta, gama, alfp, betp, gamp, s3r, vpos, mult, nmix, big_l, nLO_at,
 *   ncore, cgcoef, lmax, loctmatsize) (pygw.py:2655,
 *   for_vxcnn.f90:839-1028):
 *     mmat[(ie1 n2 + ie2) ld + col0 + im] = < u_im Y_lb,mb | psi*_ie2 psi_ie1 >
 *                                          e^{-2 pi i ql.pos_idf},
 *   ie1 < n1 = nend - nstart, ie2 < n2 = mend - mstart, im < loctmatsize:
 *   the rows of the concatenated mmat that pygw_minm_is fills after
 *   loctmatsize (col0 = 0 here).  Device arrays, PyGW's Fortran layouts:
 *   ql [3], pos (3, ndf), alfa/beta (nbmax, (lmax+1)^2, ndf) and gama
 *   (nbmax_g, nlomax, (lmax+1)^2, ndf) complex128 at k, alfp/betp/gamp
 *   likewise at k - q (mbmax, mbmax_g), s3r (n_mix, n_fnc, n_fnc, nat),
 *   cgcoef [n_cg] (cmp_all_Gaunt's table).  nlomax >= the largest nLO_at;
 *   gama/gamp may be NULL when it is 0.  FP64 (MFMA).  workspace:
 *   pygw_minm_mt_workspace bytes or NULL.
 * pygw_minc replaces fvxcn.calc_minc(kil, nst, nend, cst, cend, alfa, beta,
 *   gama, s3r, corind, ...) (pygw.py:2677, for_vxcnn.f90:1136-1273): rows
 *   (ie1 (cend - cstart) + ic) of ld >= col0 + loctmatsize; pygw_minc2
 *   replaces fvxcn.calc_minc2(kjl, cst, cend, mst, mend, alfp, betp, gamp,
 *   ...) (pygw.py:2680, for_vxcnn.f90:1275-1397): rows (ic (mend - mstart)
 *   + ie2).  corind int32 (ncg_rows, 5) Fortran (atom type, atom, core
 *   index, l, m; 0-based) on the device.  Allocate nothing.
 * ------------------------------------------------------------------------ */
int pygw_minm_mt_plan(int nat, const int* mult, const int* nmix, const int* big_l, int max_nmix, const int* nLO_at,
                      int lomax, const int* ncore, int lmax, int* plan, int* plan_ints);
int pygw_minm_mt_workspace(const int* plan_host, int n1, int n2, size_t* bytes);
int pygw_minm_mt(const int* plan_host, const int* plan_dev, const double* ql, int nstart, int nend, int mstart,
                 int mend, const void* alfa, const void* beta, const void* gama, int nbmax, int nbmax_g,
                 const void* alfp, const void* betp, const void* gamp, int mbmax, int mbmax_g, int nlomax,
                 const double* s3r, int n_mix, int n_fnc, const double* pos, const double* cgcoef, int n_cg,
                 void* mmat, long ld_mmat, int col0, void* workspace, size_t ws_bytes, void* stream);
int pygw_minc(const int* plan_host, const int* plan_dev, const double* kl, int nstart, int nend, int cstart, int cend,
              const void* alfa, const void* beta, const void* gama, int nbmax, int nbmax_g, int nlomax,
              const double* s3r, int n_mix, int n_fnc, const int* corind, int ncg_rows, const double* pos,
              const double* cgcoef, int n_cg, void* mmat, long ld_mmat, int col0, void* stream);
int pygw_minc2(const int* plan_host, const int* plan_dev, const double* kmql, int cstart, int cend, int mstart,
               int mend, const void* alfp, const void* betp, const void* gamp, int mbmax, int mbmax_g, int nlomax,
               const double* s3r, int n_mix, int n_fnc, const int* corind, int ncg_rows, const double* pos,
               const double* cgcoef, int n_cg, void* mmat, long ld_mmat, int col0, void* stream);

/* ------------------------------------------------------------------------
 * The LAPW augmentation coefficients and the band coefficients (lapw.hip).
 * pygw_lapw_plan (host only) packs what PyGW passes to
 *   lapwc.gap2_set_lapwcoef that does not depend on k (pygw.py:2571):
 *   mult [nat], lapw (nt, nat), nlo and nLO_at (lomax+1, nat) int32, umt
 *   (nat, nt, 5), abcelo (nat, lomax+1, nloat, 3), rmt [nat], pos (3, ndf),
 *   rotloc (nat, 3, 3), trotij (ndf, 3, 3), k2cartes (3, 3) float64 (host
 *   arrays, Fortran layouts) and Vol: with plan = NULL it returns the size in
 *   *plan_doubles, else fills plan.  nt <= 16; the local orbitals of each
 *   (l, atom type) must fit their lohns block (nLO_at <= nlo - 1 + lapw) and
 *   abcelo (nLO_at + 1 <= nloat); an APW l needs p = umt(:, l, 1) != 0.  The
 *   same plan is passed in host (plan_host) and device (plan_dev) memory.
 * pygw_lapw_coef replaces lapwc.gap2_set_lapwcoef(kl, indgk, iop, ifrot, nvk,
 *   gindex, abcelo, rmt, pos, mult, umt, rotloc, trotij, Vol, k2cartes,
 *   nLO_at, nlo, lapw, nLOmax) (pygw.py:2571,2595,2857;
 *   lapwcoeff.f90:3-222): alm, blm (ngk, nt^2, ndf) and clm (ngk, nLOmax,
 *   nt^2, ndf) complex128 on the device, Fortran order, every entry
 *   written.  kl [3] host; indgk [ngk] int32 (0-based rows of gindex, the
 *   first nvk plane waves), gindex (ng, 3) int32 Fortran, on the device; a
 *   row whose index is not a row of gindex is written as NaN.  clm may be
 *   NULL when the plan has no local orbitals.  Allocates nothing.
 * pygw_lapw_band replaces alfa = la.matmul(Aeig, reshape(alm, (ngk, ncol)))
 *   and its beta/gama twins (pygw.py:2577-2583, 2601-2607, 2863-2869):
 *   alfa/beta (nb, ncol), gama (nb, ncol nlomax) complex128 Fortran =
 *   A (alm, blm, clm read as (ngk, ...) Fortran), A(n, g) = Aeig[n lda + g]
 *   (conjugated with conj_a, as PyGW forms Aeigq, pygw.py:2587), ncol =
 *   nt^2 ndf.  FP64 (MFMA).  Allocates nothing.
 * ------------------------------------------------------------------------ */
int pygw_lapw_plan(int nat, const int* mult, int nt, int lomax, int nloat, const int* lapw, const int* nlo,
                   const int* nLO_at, const double* umt, const double* abcelo, const double* rmt, const double* pos,
                   const double* rotloc, const double* trotij, double vol, const double* k2cartes, double* plan,
                   int* plan_doubles);
int pygw_lapw_coef(const double* plan_host, const double* plan_dev, const double* kl, int iop, int ifrot, int nvk,
                   const int* indgk, int ngk, const int* gindex, int ng, void* alm, void* blm, void* clm,
                   void* stream);
int pygw_lapw_band(const void* Aeig, long lda, int conj_a, int nb, int ngk, const void* alm, const void* blm,
                   const void* clm, long ncol, int nlomax, void* alfa, void* beta, void* gama, void* stream);

/* ------------------------------------------------------------------------
 * The head of the dielectric matrix at q = 0 and its momentum matrix
 * elements (head.hip).
 * Common layout arguments (host memory, Fortran layouts): nat atom types
 * (<= 16) with mult [nat]; nLO_at (nrows = lomax+1 <= 17, nat) int32 (a
 * count for an l at or beyond nrows reads as 0, where PyGW's Fortran reads
 * past the array and multiplies zero gama entries); lmax = the second extent
 * of iul_ul (PyGW's in1.nt - 1): alfa/beta (nbmax, (lmax+1)^2, ndf) and gama
 * (nbmax_g, nlomax, (lmax+1)^2, ndf) complex128 on the device, gama NULL
 * when nlomax = 0.
 * pygw_mmat_vv replaces calc_head's double loop over bands (pygw.py:
 *   2870-2909): f_q_0.calcmmatvv_mt2 / calcmmatvv_mt (for_q_0.f90:10-458)
 *   and f_q_0.calcmmatvv_is (for_q_0.f90:460-507) in both orders, averaged:
 *   mmatvv[(ie2 - mst) (nend - nst) + ie1 - nst][c] complex128.  vv_tables:
 *   a host array of 9 device pointers (iul_ul, iul_udl, iudl_ul, iudl_udl
 *   (2, lmax, nat); iulol_ul, iulol_udl, iul_ulol, iudl_ulol (2, nlo, lox,
 *   nat); iulol_ulol (2, nlo, nlo, lox, nat); float64 Fortran; the last five
 *   may be NULL when nlo = 0).  Interstitial part (nvk > 0): Aeig [nbmax][lda]
 *   complex128 row-major, indgk [>= nvk] int32 (0-based rows of gindex (ng, 3)
 *   int32 Fortran), ipwint [ng] complex128, i_g0 (2 n_g0[0]+1, 2 n_g0[1]+1,
 *   2 n_g0[2]+1) int32 Fortran on the device; kl [3], n_g0 [3], k2cartes (3, 3)
 *   in host memory.  FP64 (the products on MFMA); workspace:
 *   pygw_mmat_vv_workspace bytes (ndf = sum of mult).
 * pygw_mmat_cv replaces f_q_0.calcmmatcv over the atoms with cores (pygw.py:
 *   2911-2932, for_q_0.f90:510-683): mmatcv[(ie2 - mst) ncg + icg][c].
 *   cv_tables: 6 device pointers (iul_ucl, iudl_ucl, iucl_ul, iucl_udl (2,
 *   nct); iulol_ucl, iucl_ulol (2, nloat, nct)), the core states of atom type
 *   iat from core_off[iat] (host) on; corind (ncg, 5) int32 Fortran on the
 *   device; a core state with l + 1 > lmax or outside the tables reads as
 *   NaN.  exact = 0 reproduces the reference (local orbitals at l - 1 times
 *   iudl_ucl in one term, for_q_0.f90:648), 1 takes iulol_ucl there.
 * pygw_calc_head replaces calc_head's sums (pygw.py:2940-3025): head[nom]
 *   complex128 = -sum_irk 4 pi fspin weight[irk]/vol sum_pairs |p|^2/3/dE^2
 *   kcw(pair, k_ind[irk], iom) (valence pairs with |dE| < 1e-5 dropped), plus
 *   wpl2/(omega + eta_head)^2 (projected on dUl (nom_w, nom) row-major when
 *   given) for metallic with iop_drude = 1, wpl2 from kwfer [nirk][ldE].
 *   mmatvv [nirk][nbmaxpol-numin][nomx+1][3], mmatcv [nirk][..][ncg][3],
 *   Ebnd [nirk][ldE], eig_core [ncg], kcw (ncg_p+nomx+1, nbmaxpol-numin,
 *   nk_kcw, nom) Fortran, k_ind [nirk] int32, weight [nirk], all on the
 *   device; nom <= 4096; workspace: pygw_calc_head_workspace bytes.
 * pygw_wing_pm replaces calc_eps's pm of the q = 0 wing (pygw.py:3122-3141):
 *   pm[k][(io (nbmaxpol - numin) + ie2] for the k-points whose irreducible
 *   index is kii[k] (< nirk, the first extent of mmatvv; a k whose index is
 *   not reads as NaN), the input of pygw_chi0_wings.  In pygw_calc_head a
 *   k_ind outside kcw's k-points makes the head NaN.  Allocates nothing.
 * ------------------------------------------------------------------------ */
int pygw_mmat_vv_workspace(int ndf, int lmax, int nlomax, int nvk, int nst, int nend, int mst, int mend,
                           size_t* bytes);
int pygw_mmat_vv(int nat, const int* mult, const int* nLO_at, int nrows, int lmax, const void* alfa, const void* beta,
                 const void* gama, int nbmax, int nbmax_g, int nlomax, const double* const* vv_tables, int nlo,
                 int lox, const void* Aeig, long lda, const double* kl, const int* indgk, int nvk, const void* ipwint,
                 const int* gindex, int ng, const int* i_g0, const int* n_g0, const double* k2cartes, int nst,
                 int nend, int mst, int mend, void* mmatvv, void* workspace, size_t ws_bytes, void* stream);
int pygw_mmat_cv(int nat, const int* mult, const int* nLO_at, int nrows, int lmax, const int* core_off,
                 const void* alfa, const void* beta, const void* gama, int nbmax, int nbmax_g, int nlomax,
                 const double* const* cv_tables, int nloat, int nct, const int* corind, int ncg, int mst, int mend,
                 int exact, void* mmatcv, void* stream);
int pygw_calc_head_workspace(int nirk, int n_o, int n_e, size_t* bytes);
int pygw_calc_head(const void* mmatvv, const void* mmatcv, const double* Ebnd, int ldE, const double* eig_core,
                   int ncg, const double* kcw, int nk_kcw, const int* k_ind, const double* weight, int nirk, int nomx,
                   int numin, int nbmaxpol, int ncg_p, int nom, double vol, int metallic, const double* kwfer,
                   int iop_drude, double eta_head, const double* omega, const double* dUl, int nom_w, void* head,
                   void* workspace, size_t ws_bytes, void* stream);
int pygw_wing_pm(const void* mmatvv, const void* mmatcv, const double* Ebnd, int ldE, const double* eig_core, int ncg,
                 int nirk, const int* kii, int nk, int nomx, int numin, int nbmaxpol, int ncg_p, double vol, void* pm,
                 void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PYGW_HIP_H */
