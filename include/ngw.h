/*
 * ngw.h -- C ABI of libngw, the gfx950 (MI355X) ISDF-Lanczos GW hot path.
 *
 * Drop-in boundary for NanoGW's Lanczos self-energy path.  Every entry point
 * names the NanoGW routine (file:line) it replaces.  All arrays are Fortran
 * column-major with 0-based C pointers; complex arrays are interleaved
 * (re, im) pairs of double, i.e. Fortran complex(dp).  Energies are Rydberg
 * (NanoGW's internal unit, e^2 = 2) unless an argument says eV.
 *
 * Pair ordering of transition-space vectors (length ncv = nval * ncond):
 *   icv = iv + nval * ic   (0-based iv < nval, ic < ncond; band of ic is
 *   nval + ic), the conduction-outer / valence-inner order NanoGW builds in
 *   shared/setup_g.F90z:154-168.
 *
 * Host pointers: every entry point below takes HOST arrays, like the Fortran
 * callers it replaces, except the *_dev entry points, which take device
 * pointers and run on the context's stream.  The operator (Psi at the ISDF
 * points, M, energies) stays resident on the GPU inside the context between
 * calls.
 *
 * Errors: functions return NGW_OK (0) or a negative status; ngw_last_error()
 * gives the message.  NanoGW reacts to errors with die() (shared/die.F90:10);
 * the Fortran shim maps a non-zero status to die().
 * Threading: one context is used by one host thread at a time.
 */
#ifndef NGW_H
#define NGW_H

#ifdef __cplusplus
extern "C" {
#endif

#define NGW_OK 0
#define NGW_ERR_ARG (-1)      /* invalid argument or size */
#define NGW_ERR_STATE (-2)    /* operator not uploaded / inconsistent sizes */
#define NGW_ERR_HIP (-3)      /* HIP runtime or hipFFT failure */
#define NGW_ERR_NOMEM (-4)    /* device allocation failed */
#define NGW_ERR_RANGE (-5)    /* polyfit_sqrt: xmin <= 1e-4 (NanoGW dies there) */

#define NGW_PREC_F64 0   /* all GEMMs FP64 (the parity path) */
#define NGW_PREC_F32 1   /* C x, M t and C^T w GEMMs in FP32, FP32 accumulate */
#define NGW_PREC_BF16 2  /* C x and C^T w in BF16 with per-band Psi scaling, FP32 accumulate;
                            M t in FP32 */
#define NGW_PREC_BF16X3 3 /* as BF16 with each operand split into BF16 head + tail, three
                             MFMAs per product (about 16 mantissa bits) */
#define NGW_PREC_F16 4   /* C x and C^T w in FP16 (11 mantissa bits) with exact power-of-two
                            scales per band and per vector, FP32 accumulate; M t in FP32 */

typedef struct ngw_context ngw_context;

/* Library version string and the last error message of this thread. */
const char* ngw_version(void);
const char* ngw_last_error(void);

/* Context on HIP device `device` with its own stream. */
int ngw_create(ngw_context** ctx, int device);
int ngw_destroy(ngw_context* ctx);
/* Wait for all work queued on the context's stream. */
int ngw_synchronize(ngw_context* ctx);

/*
 * Upload the ISDF operator, replacing the fields of type(ISDF) the Lanczos
 * path reads (shared/typedefs.F90:24-72):
 *   psi_intp(ld_psi, nband)  Psi_intp_loc(:, 1:nband, 1, 1): orbitals at the
 *                            n_intp ISDF points, bands 1..nval occupied
 *   mmtrx(ld_m, n_intp)      Mmtrx_loc(:, :, 1, 1, 1, 1, 1): Coulomb kernel
 *   e1(nband)                kpt%wfn(1,1)%e1 (Ry)
 * All nval x (nband - nval) pairs form the transition space.
 */
int ngw_set_isdf(ngw_context* ctx, int n_intp, int nband, int nval, const double* psi_intp,
                 int ld_psi, const double* mmtrx, int ld_m, const double* e1);

/* GEMM precision of the operator (NGW_PREC_*); the Lanczos alpha/beta, the
 * continued fraction, Sigma_c assembly and the QP Newton solve stay FP64.
 * In the low-precision modes p(H) is evaluated as the same polynomial
 * re-expressed in Chebyshev form on [e_min, e_max] of the driver's fit
 * (calculate_sigma_lanczos_BLOCK.F90z:166-171) with the Clenshaw recurrence,
 * which does not amplify rounding the way monomial Horner does. */
int ngw_set_precision(ngw_context* ctx, int precision);

#define NGW_REORTH_NONE 0 /* the reference's plain three-term recurrence
                             (shared/lanczos_spectra_BLOCK.F90:79-108) */
#define NGW_REORTH_FULL 1 /* full reorthogonalisation: after w -= alpha q_k, two
                             classical Gram-Schmidt passes of w against q_1..q_k */

/* Reorthogonalisation of the Lanczos recurrences of ngw_lanczos_spectra_isdf_block
 * and ngw_sigma_lanczos(_bands) (NGW_REORTH_*; default NONE, which is NanoGW's
 * algorithm and the parity path).  No counterpart in NanoGW.  FULL keeps every
 * Lanczos vector on the device, 8 * niter * ncv bytes per column; a run that
 * does not fit returns NGW_ERR_NOMEM.  All arithmetic FP64 in every precision
 * mode.  The batched engine (ngw_batch_*) runs the plain recurrence only. */
int ngw_set_reorthogonalization(ngw_context* ctx, int mode);

/*
 * Upload the operator in the explicit form of NanoGW's single-vector
 * routines, matvec_isdf (shared/apply_matvec.f90:25), polynomial_matvec_isdf
 * (:142) and lanczos_spectra_isdf (shared/lanczos_spectra.f90:7), which take
 * the pair matrix and the energy factors as arguments:
 *   cmtrx(ld_c, ncv)      tmpCmtrx(n_intp_r, ncv_loc), C(mu, icv) =
 *                         Psi_v(mu) Psi_c(mu) for any pair list
 *   mmtrx(ld_m, n_intp)   isdf_in%Mmtrx(:, :, 1, 1, 1, 1, 1)
 *   sqrtr(ncv)            sqrtR, (e_c - e_v)^1/2 > 0
 * After it the operator entries (ngw_matvec_isdf_block,
 * ngw_polynomial_matvec_isdf_block, ngw_lanczos_spectra_isdf_block and the
 * _dev apply) compute the same y = R^2 x + 4 R^1/2 C^T M C R^1/2 x (TDA:
 * y = R x + 2 C^T M C x) with the three products as this library's FP64
 * GEMV / GEMM kernels.  FP64 only; the Sigma_c / Sigma_x entries and
 * ngw_autotune need the factored operator and fail with NGW_ERR_STATE.
 * ngw_set_isdf switches back.
 */
int ngw_set_explicit_c(ngw_context* ctx, int n_intp, int ncv, const double* cmtrx, int ld_c,
                       const double* mmtrx, int ld_m, const double* sqrtr);

/*
 * Irrep-blocked operator of the lowcomsym routines, replaces the operator
 * data of matvec_isdf_lowcomsym / polynomial_matvec_isdf_lowcomsym /
 * lanczos_spectra_isdf_lowcomsym (shared/apply_matvec_lowcomsym.F90:31,
 * :419; shared/lanczos_spectra_lowcomsym.F90:7): isdf_in%PsiV_intp_bl /
 * PsiC_intp_bl gathered through lvrep / lcrep, Mmtrx_loc(:, :, ..., nmrep)
 * and sqrtR.  Block k (k = 0 .. nblock-1) is one valence irrep and the
 * conduction irrep whose product with it is nmrep (:158-161): its nv_blk[k]
 * valence states are columns sum_{k'<k} nv_blk[k'] .. of psiv(ld_v, *), its
 * nc_blk[k] conduction states the matching columns of psic(ld_c, *).  Pairs
 * are numbered block after block, valence state outer and conduction state
 * inner (the reference's cvstart order, :159-247), ncv = sum nv_blk nc_blk;
 * sqrtr(ncv) in that order.  The operator entries then apply
 * y = R^2 x + 4 R^1/2 C^T M C R^1/2 x (TDA: y = R x + 2 C^T M C x) in
 * factored form, block by block, without forming C.  FP64 only; the
 * Sigma_c / Sigma_x entries and ngw_autotune fail with NGW_ERR_STATE.
 * ngw_set_isdf or ngw_set_explicit_c switches to another form.
 */
int ngw_set_isdf_blocks(ngw_context* ctx, int n_intp, int nblock, const int* nv_blk, const int* nc_blk,
                        const double* psiv, int ld_v, const double* psic, int ld_c,
                        const double* mmtrx, int ld_m, const double* sqrtr);

/*
 * Operator apply on blksz vectors, replaces matvec_isdf_BLOCK
 * (shared/apply_matvec_BLOCK.F90:25) and, with tamm_d = 1, the TDA branch of
 * matvec_isdf_lowcomsym (shared/apply_matvec_lowcomsym.F90:31, :396-403):
 *   full: hvec = R^2 vec + 4 R^1/2 C^T M C R^1/2 vec
 *   TDA : hvec = R vec + 2 C^T M C vec
 * vec, hvec: (ncv, blksz).  ncv must equal nval*ncond (the columns of C
 * after ngw_set_explicit_c).
 */
int ngw_matvec_isdf_block(ngw_context* ctx, const double* vec, double* hvec, int ncv,
                          int blksz, int tamm_d);

/*
 * Horner p(H) vec, replaces polynomial_matvec_isdf_BLOCK
 * (shared/apply_matvec_BLOCK.F90:210): polynomial(1..npoly+1) holds
 * a_n .. a_0 (highest order first); npoly matvecs.
 */
int ngw_polynomial_matvec_isdf_block(ngw_context* ctx, const double* vec, double* hvec,
                                     int ncv, const double* polynomial, int npoly, int blksz,
                                     int tamm_d);

/*
 * blksz independent Lanczos recurrences on H_aux = p(H) (no
 * reorthogonalisation) plus the continued fraction
 * spectra(iz, j) = <v0_j| (zz(iz, j) - H_aux)^-1 |v0_j> / |v0_j|^2 ... as in
 * lanczos_spectra_isdf_BLOCK (shared/lanczos_spectra_BLOCK.F90:7):
 *   v0(ncv, blksz), v0_norm(blksz), zz(nz, blksz) complex,
 *   spectra(nz, blksz) complex (output).
 * a_elmt, b_elmt (niter, blksz) receive the recurrence coefficients when not
 * NULL.  After ngw_set_explicit_c, blksz = 1 is lanczos_spectra_isdf
 * (shared/lanczos_spectra.f90:7) but for its breakdown exit at |b_k| < 1e-8
 * (:58), which the caller applies to the returned coefficients.
 */
int ngw_lanczos_spectra_isdf_block(ngw_context* ctx, const double* v0, const double* v0_norm,
                                   int ncv, const double* zz, int nz, int niter,
                                   const double* polynomial, int npoly, double* spectra,
                                   int blksz, int tamm_d, double* a_elmt, double* b_elmt);

/*
 * Sigma_c(omega) of the Lanczos GW drivers for nqp QP states, replaces the
 * GW body of Zcalculate_sigma_lanczos_BLOCK
 * (SIGMA/calculate_sigma_lanczos_BLOCK.F90z:201-444; TDA increments from
 * SIGMA/calculate_sigma_lanczos_lowcomsym.F90z:712-745):
 *   qp_band(nqp)   0-based band index m of each QP state
 *   nen, deltae    energy_data_points and energy_range (Ry); the grid is
 *                  z_j = e1(m) - deltae/2 + j deltae/(nen-1)
 *   ecuts          dynamic_energy_resolution gamma (Ry)
 *   niter, polynomial(npoly+1)  lanczos_niter and the sqrt polynomial
 *   sigma_c(nen, nqp) complex (Ry, output)
 * Every (QP state, band) Lanczos run is done on the device in batched form;
 * max_cols (0 = automatic) bounds how many runs share one batch.
 */
int ngw_sigma_lanczos(ngw_context* ctx, int nqp, const int* qp_band, int nen, double deltae,
                      double ecuts, int niter, const double* polynomial, int npoly, int tamm_d,
                      int max_cols, double* sigma_c);

/*
 * The part of ngw_sigma_lanczos contributed by the Green's-function bands
 * band_lo <= n < band_hi (0-based): the band loop of
 * SIGMA/calculate_sigma_lanczos_BLOCK.F90z:201-416 restricted to one
 * contiguous range.  Sigma_c is a plain sum over bands, so ranges computed
 * on different devices add up to the full result; the unit of work a
 * multi-device driver would hand to each GPU.
 */
int ngw_sigma_lanczos_bands(ngw_context* ctx, int nqp, const int* qp_band, int band_lo, int band_hi, int nen,
                            double deltae, double ecuts, int niter, const double* polynomial, int npoly,
                            int tamm_d, int max_cols, double* sigma_c);

/*
 * Bare exchange of QP states in the ISDF representation, replaces the
 * diagonal matrix elements of Zfock_exchange (SIGMA/fock_exchange.F90z:12-229)
 * with the operator's own Coulomb kernel:
 *   sigx(q) = - sum_{v < nval} sum_{mu,nu} c(mu) M(mu,nu) c(nu),
 *   c(mu) = Psi(mu, qp_band(q)) Psi(mu, v)          (Ry)
 * qp_band(nqp) 0-based bands; sigx(nqp) output.
 */
int ngw_sigma_x_isdf(ngw_context* ctx, int nqp, const int* qp_band, double* sigx);

/*
 * QP energies by Newton-Raphson on the cubic spline of
 * f(E) = E_dft + Sigma_x + Sigma_c(E) - V_xc - E (eV), replaces the diagonal
 * block of printsigma (SIGMA/printsigma.F90:97-128) with newtonr
 * (SIGMA/newtonr.f90:18), spline_evaluate, spline and splint; run on the GPU
 * in FP64, one state per thread.
 *   nqp states, nen grid points, deltae_ev = energy_range in eV
 *   edft(nqp) = e0*ryd, en0(nqp) = e1*ryd, sigx(nqp), vxc(nqp) real, eV
 *   sigma_c(nen, nqp) complex, eV
 * Outputs (nqp): eqp, zfac, sigma0 (complex), imsig, ierr (0 = converged).
 */
int ngw_qp_newton(ngw_context* ctx, int nqp, int nen, double deltae_ev, const double* edft,
                  const double* en0, const double* sigx, const double* vxc,
                  const double* sigma_c, double* eqp, double* zfac, double* sigma0,
                  double* imsig, int* ierr);

/* Largest nqp * nen of ngw_qp_newton_dev_noalloc (its work arrays take
 * 5 * 8 bytes per point: 2.5 MiB per GPU). */
#define NGW_QP_NOALLOC_MAX_POINTS 65536

/*
 * ngw_qp_newton on device-resident data, allocating nothing: no hipMalloc,
 * hipFree or allocating library call; the spline work arrays live in a
 * fixed per-GPU device array (NGW_QP_NOALLOC_MAX_POINTS points) that the
 * library keeps until the process ends.  No counterpart in NanoGW: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation).
 *   d_in  device, 4 * nqp + 2 * nen * nqp doubles: (edft, en0, sigx, vxc) of
 *         each state (eV), then sigma_c(nen, nqp) complex (eV)
 *   d_out device, 6 * nqp doubles per state: eqp, zfac, Re sigma0,
 *         Im sigma0, imsig, ierr (as a double; 0 = converged)
 * Queued on the context's stream; the caller synchronises (ngw_synchronize).
 * nqp * nen above the limit is NGW_ERR_ARG.  Every call on one GPU uses the
 * same work arrays: calls that overlap on one GPU (several contexts or
 * threads) race.
 */
int ngw_qp_newton_dev_noalloc(ngw_context* ctx, int nqp, int nen, double deltae_ev, const double* d_in,
                              double* d_out);

/*
 * Least-squares monomial fit of sqrt(x) on [xmin, xmax] through nx nodes,
 * replaces polyfit_sqrt (shared/polyfit.f90:1) and, with chebyshev = 1,
 * polyfitcheb_sqrt (shared/polyfitcheb.f90:1); beta(npoly+1) is returned
 * lowest order first (beta(1) = a_0) as in NanoGW; the Lanczos entry points
 * take the reversed order.  Host-side, FP64.
 */
int ngw_polyfit_sqrt(double xmin, double xmax, int npoly, int nx, int chebyshev, double* beta,
                     double* fit_error);

/*
 * ISDF Coulomb kernel M = (ntrans / h^3) zeta^T V zeta, replaces the V step
 * of isdf_parallel_sym_UltraLowMem (shared/isdf_parallel_sym_UltraLowMem.F90:
 * 686-875) with the isolated-system Poisson solve Zpoisson
 * (shared/fft_module.F90z:122-177) and the truncated kernel
 * v(G) = 8 pi (1 - cos(|G| R_c)) / |G|^2, v(0) = 8 pi R_c^2 / 2
 * (Zcreate_coul_0D, shared/fft_module.F90z:628-691):
 *   zeta(ld_zeta, n_intp)  interpolation vectors on the ngrid points
 *   grid(3, ngrid)         integer grid coordinates in [-nfft/2, nfft/2)
 *   nfft[3]                FFT box; h grid step; trunc = R_c
 *   mmtrx(ld_m, n_intp)    output
 * Batched R2C/C2R FFTs (hipFFT) with hand-written gather / kernel-multiply /
 * scatter kernels and an FP64 MFMA GEMM for zeta^T (V zeta).
 */
int ngw_isdf_coulomb_kernel(ngw_context* ctx, int ngrid, int n_intp, const double* zeta,
                            long ld_zeta, const int* grid, const int* nfft, double h,
                            double trunc, double ntrans, double* mmtrx, int ld_m);

/*
 * Same V build with zeta already on the device (d_zeta, ld_zeta), as left by
 * ngw_isdf_zeta_dev: the psi -> zeta -> M chain stays device-resident.
 */
int ngw_isdf_coulomb_kernel_dev(ngw_context* ctx, int ngrid, int n_intp, const double* d_zeta,
                                long ld_zeta, const int* grid, const int* nfft, double h,
                                double trunc, double ntrans, double* mmtrx, int ld_m);

/*
 * ISDF interpolation vectors, replaces the zeta solve of
 * isdf_parallel_sym_UltraLowMem (shared/isdf_parallel_sym_UltraLowMem.F90:
 * 448-642: P, Q GEMMs, A = P o Q at the points, B = P o Q on the grid,
 * pdgesv) on the device:
 *   psi(ld_psi, nband)   orbitals on the ngrid points, occupied bands first
 *   intp(n_intp)         0-based grid indices of the interpolation points
 *   A(mu,nu) = [sum_{v<nval} psi(mu,v) psi(nu,v)] [sum_n psi(mu,n) psi(nu,n)]
 *              (+ ridge tr(A)/n_intp on the diagonal when ridge > 0)
 *   B(r,nu)  = [sum_{v<nval} psi(r,v) psi(nu,v)] [sum_n psi(r,n) psi(nu,n)]
 *   zeta(ld_zeta, n_intp) = B A^-1
 * A^-1 by Cholesky (LU when A is not numerically positive definite:
 * *used_lu = 1).  The _dev form takes and returns device arrays.
 */
int ngw_isdf_zeta(ngw_context* ctx, int ngrid, int nband, int nval, const double* psi, long ld_psi,
                  int n_intp, const int* intp, double ridge, double* zeta, long ld_zeta, int* used_lu);
int ngw_isdf_zeta_dev(ngw_context* ctx, int ngrid, int nband, int nval, const double* d_psi, long ld_psi,
                      int n_intp, const int* intp, double ridge, double* d_zeta, long ld_zeta, int* used_lu);

/*
 * ISDF interpolation points by rho-weighted centroidal Voronoi tessellation,
 * replaces cvt (shared/cvt.F90:8-195): Lloyd iterations of nearest-centroid
 * assignment and rho-weighted centroids (on the GPU), each centroid snapped
 * in index order to the nearest grid point not already taken, until the
 * summed displacement is below tol (bohr) or max_iter iterations.
 *   grid(3, ngrid)    integer lattice coordinates (bohr = h x coordinate)
 *   rho(ngrid)        weights; init(3, n_intp) starting centroids (bohr)
 *   intp(n_intp)      output: 0-based grid indices; *iters iterations run
 */
int ngw_cvt_points(ngw_context* ctx, int ngrid, const int* grid, const double* rho, int n_intp,
                   const double* init, double h, double tol, int max_iter, int* intp, int* iters);

/*
 * Device-pointer operator apply for callers that keep vectors on the GPU
 * (the serving layer): same as ngw_matvec_isdf_block with d_vec, d_hvec
 * device pointers of (ncv, blksz) with leading dimension ld.
 */
int ngw_matvec_isdf_block_dev(ngw_context* ctx, const double* d_vec, double* d_hvec, long ld,
                              int blksz, int tamm_d);

/*
 * FP64 GEMM C(m,n) = A(m,k) B(k,n) on device pointers (column-major), the
 * MFMA kernel the operator uses; exported for the library's own tests of
 * the gfx950 matrix-core lane layout.
 */
int ngw_gemm_f64_dev(ngw_context* ctx, int m, int n, int k, const double* d_a, long lda,
                     const double* d_b, long ldb, double* d_c, long ldc);

/*
 * Tile-shape autotuner of the low-precision operator GEMMs (no NanoGW
 * counterpart: NanoGW calls vendor BLAS).  For the context's current
 * precision and operator sizes, applies the operator to nvec vectors with
 * each of the NGW_NTILE_CONFIGS {BM, BN, slabs-in-flight, XCD M-groups}
 * configurations of K1 (C x) and K3 (C^T w), reps timed applications each,
 * keeps the fastest of each for later calls on this context (and batches
 * created from it) and returns the chosen indices.  A no-op in FP64 mode.
 * ngw_set_tile_config sets the indices (0 .. NGW_NTILE_CONFIGS - 1)
 * directly; index 0 is the default.
 */
#define NGW_NTILE_CONFIGS 6
int ngw_autotune(ngw_context* ctx, int nvec, int reps, int* k1_cfg, int* k3_cfg);
int ngw_set_tile_config(ngw_context* ctx, int k1_cfg, int k3_cfg);

/*
 * Per-launch kernel timing with HIP events on the context's stream (no
 * NanoGW counterpart; NanoGW's timacc slots 68-78 time the same stages on the
 * host, shared/apply_matvec_BLOCK.F90:52-206).  Kernel classes:
 * 0 K1 (C x GEMM), 1 K2 (M t GEMM), 2 K3 (C^T w GEMM + epilogue),
 * 3 Lanczos alpha/beta, 4 continued fraction, 5 P-vector build,
 * 6 Sigma_c assembly, 7 low-precision operand conversion (x' and W to
 * FP32/BF16, slot reduction).  ngw_profile_read synchronizes and returns the
 * summed milliseconds and launch count of one class since the last reset.
 */
#define NGW_KC_COUNT 8
int ngw_profile_enable(ngw_context* ctx, int on);
int ngw_profile_reset(ngw_context* ctx);
int ngw_profile_read(ngw_context* ctx, int kernel_class, double* total_ms, long* launches);

/*
 * C = A B through the low-precision MFMA core (kind 0 FP32, 1 BF16,
 * 2 BF16x3, 3 FP16 unscaled), FP64 device arrays in and out; exported for the library's own
 * tests of the matrix-core lane layouts.
 */
int ngw_gemm_lowp_dev(ngw_context* ctx, int kind, int m, int n, int k, const double* d_a, long lda,
                      const double* d_b, long ldb, double* d_c, long ldc);

/*
 * Batched and continuously served GW (no single NanoGW counterpart: it runs
 * the per-request work of Zcalculate_sigma_lanczos_BLOCK,
 * SIGMA/calculate_sigma_lanczos_BLOCK.F90z:201-444, for many molecules at
 * once).  A batch uses its context's device, stream and precision.
 *   add_system     upload one molecule's ISDF operator (kept resident)
 *   request_bytes  HBM a request would allocate (for admission control)
 *   submit         one GW request: nqp QP states of a system; builds the P
 *                  vectors and the Lanczos start on the device
 *   step           advance every unfinished request by nsteps Lanczos steps
 *                  in grouped launches (use_graph = 1: one hipGraph per step,
 *                  re-captured when requests join or leave); requests that
 *                  reach niter get their continued fraction and Sigma_c
 *   result         Sigma_c(nen, nqp) complex (Ry) once done
 *   hold           hold = 1 keeps an unfinished request out of the steps (its
 *                  Lanczos state stays resident), hold = 0 lets it step again:
 *                  the serving loop's lever for giving short requests the GPU
 * A request's result does not depend on the other requests in the batch, nor
 * on when it was held.
 */
typedef struct ngw_batch ngw_batch;
int ngw_batch_create(ngw_context* ctx, ngw_batch** batch);
int ngw_batch_destroy(ngw_batch* batch);
int ngw_batch_add_system(ngw_batch* batch, int n_intp, int nband, int nval, const double* psi_intp,
                         int ld_psi, const double* mmtrx, int ld_m, const double* e1, int* sys_id);
int ngw_batch_remove_system(ngw_batch* batch, int sys_id);
int ngw_batch_request_bytes(ngw_batch* batch, int sys_id, int nqp, int niter, int nen, long* bytes);
int ngw_batch_submit(ngw_batch* batch, int sys_id, int nqp, const int* qp_band, int nen, double deltae,
                     double ecuts, int niter, const double* polynomial, int npoly, int tamm_d,
                     int* req_id);
int ngw_batch_step(ngw_batch* batch, int nsteps, int use_graph, int* n_finished);
int ngw_batch_active(ngw_batch* batch, int* n_active);
int ngw_batch_result(ngw_batch* batch, int req_id, double* sigma_c, int* done);
int ngw_batch_release(ngw_batch* batch, int req_id);
int ngw_batch_hold(ngw_batch* batch, int req_id, int hold);
int ngw_batch_stats(ngw_batch* batch, long* steps_run, long* graph_captures);

#ifdef __cplusplus
}
#endif

#endif /* NGW_H */
