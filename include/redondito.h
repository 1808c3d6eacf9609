/*
 * redondito.h -- C-ABI of the MI355X-native pseudospectral transform path.
 *
 * Every function returns an int status (RD_OK = 0, negative on error) and
 * takes plain pointers and sizes.  Field arrays are DEVICE pointers (HBM);
 * calls are ordered on the plan's HIP stream (rd_plan_set_stream) and return
 * without waiting unless the name says otherwise; rd_plan_sync waits.
 *
 * Layouts are the reference's (one rank holds everything):
 *   spectral a(lm, k): complex FP64 interleaved (re, im), lm fastest, lm in
 *     the SHTns m-major order LM(l,m) = m(2 lmax + 1 - m)/2 + l, with a
 *     plane stride of nlm complex values;
 *   physical f(theta, phi, k): real FP64, theta fastest, nlat*nphi per plane.
 * Each entry point names the reference routine it replaces (paths relative
 * to the reference tree).
 */
#ifndef REDONDITO_H
#define REDONDITO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RD_OK 0
#define RD_ERR_ARG (-1)     /* bad argument or size                        */
#define RD_ERR_HIP (-2)     /* a HIP runtime call failed                   */
#define RD_ERR_TABLE (-3)   /* FC-Gram table missing or of the wrong size  */
#define RD_ERR_NOMEM (-4)   /* device allocation failed                    */
#define RD_ERR_LU (-5)      /* singular boundary matrix in setup           */
#define RD_ERR_STATE (-6)   /* call needs a setup that was not done        */

/* Precision forms of rd_nonlinear_uxw. */
#define RD_PREC_FP64 0      /* plain FP64 path (MFMA F64 Legendre)          */
#define RD_PREC_FP32 1      /* Legendre contractions on FP32 MFMA           */
#define RD_PREC_BF16X3 2    /* BF16 3-term split, FP32 accumulation         */
#define RD_PREC_OZAKI 3     /* FP64 operands as int8 digit slices, exact    */
                            /* int32 accumulation, FP64 recombination       */

typedef struct rd_plan rd_plan;

/* Human-readable text for a status code. */
const char *rd_error_string(int code);

/* Reads A{C}-{d}.dat and Q{d}.dat (Fortran stream FP64, column major) from
 * tdir and forms dir = A.Q^T (C x d, column major) on the host.
 * Replaces load_dirichlet_tables, src/sphp/fcgram_mod.f90:186-263. */
int rd_load_dirichlet_tables(const char *tdir, int C, int d, double *dir);

/* Neumann reconstructor of order ord (1: first, 2: second normal derivative)
 * for matching points at spacing dx: reads Q{d}.dat and Q{ord}n{d}.dat (a
 * leading FP64 dxp, then Qn, d x d column major) from tdir and forms
 * neu = Q(d,:) Qn^T with neu(d) scaled by (dx/dxp)^ord, d values, on the host.
 * An undefined order prints a warning and returns RD_OK without writing.
 * Replaces load_neumann_tables, src/sphp/fcgram_mod.f90:267-371. */
int rd_load_neumann_tables(const char *tdir, int d, int ord, double dx, double *neu);

/* Creates the plan: Gauss-Legendre grid, associated-Legendre tables (FP64,
 * built on the device), phi and radial FFT twiddles, the FC-Gram operator
 * dir (C x d column major, may be NULL when C = 0), the radial grid
 * z_k = rint + k dz and kz, and the device workspace.  nz counts the C
 * continuation points as the reference does (nr = nz - C physical shells).
 * Replaces fcgram_create_plan (src/sphp/fcgram_mod.f90:49-150) with the two
 * fftp3d_create_plan calls (src/sphp/sphp_sca.F90:70-144; device part
 * src/sphp-gpu/sphp_sca.F90:478-625) and the grid part of box_init
 * (src/pseudo/pseudospec_mod.f90:126-236). */
int rd_plan_create(rd_plan **plan, int lmax, int nlat, int nphi, int nz, int C, int d,
                   double dz, double rint, const double *dir);

/* Replaces fcgram_destroy_plan (src/sphp/fcgram_mod.f90:153-182). */
int rd_plan_destroy(rd_plan *plan);

/* Orders the plan's work on an existing HIP stream; NULL selects the default
 * (null) stream.  A new plan runs on a non-blocking stream of its own. */
int rd_plan_set_stream(rd_plan *plan, void *hip_stream);

/* Digits per operand of the plan's later RD_PREC_OZAKI calls: 7 (the
 * default), 6 or 5.  The tables are sliced to 7 digits once; a call with D
 * digits uses their first D digit planes and the digit products with
 * p + q <= D - 1 (28, 21 or 15 int8 GEMMs), each element exact to
 * 2^-(7 D + 1) of its row or column maximum.  Measured at lmax = 255,
 * nr = 128 (joint relative L2 against FP64): u x omega 1.4e-12 with 7,
 * 2.0e-10 with 6, 2.2e-8 with 5; the u x B part of the lmax = 511 coupled
 * step needs 7.
 * RD_ERR_ARG for other counts.  No reference routine: the reference has no
 * reduced-precision form. */
int rd_plan_set_ozaki_digits(rd_plan *plan, int digits);

/* Waits for all work queued on the plan's stream (the reference's
 * device_sync, src/memmgt/gdevice_mod.f90). */
int rd_plan_sync(rd_plan *plan);

/* info[0..8] = lmax, nlat, nphi, nlm, nz, C, d, nr, 0. */
int rd_plan_info(const rd_plan *plan, int *info);

/* Per-stage HIP-event timing of rd_nonlinear_uxw (stages: radial step,
 * synthesis of u and omega (one launch in the FP64 and FP32 forms),
 * phi/cross step, analysis), recorded on
 * the plan's stream without host synchronisation and read back by
 * rd_plan_stage_times (which waits).  No counterpart in the reference, whose
 * gtimer handles (src/utils/gtimer_mod.f90) time host wall clock only.
 * rd_plan_stage_times returns the number of stages on success. */
int rd_plan_profile(rd_plan *plan, int enable);
int rd_plan_stage_times(rd_plan *plan, double *ms, long *count, int nmax);

/* Host copies of the grid: cos/sin(theta), Gauss weights (nlat), z and kz
 * (nz), l and m per lm (nlm).  Any pointer may be NULL.  Replaces the grid
 * and wavenumber construction of box_init
 * (src/pseudo/pseudospec_mod.f90:168-227). */
int rd_plan_grid(const rd_plan *plan, double *ct, double *st, double *wt, double *z,
                 double *kz, int *ell, int *emm);

/* ---- 2D transforms, plane by plane (nplanes <= nz) ------------------- */

/* in(nlat,nphi,nplanes) -> out(nlm,nplanes).  The input is left intact.
 * Replaces fftp2d_real_to_complex_xy, src/sphp/sphp_sca.F90:273-329. */
int rd_sht_analysis(rd_plan *plan, int nplanes, const double *in, double *out);

/* in(nlm,nplanes) -> out(nlat,nphi,nplanes); Im of m = 0 is ignored.
 * Replaces fftp2d_complex_to_real_xy, src/sphp/sphp_sca.F90:436-492. */
int rd_sht_synthesis(rd_plan *plan, int nplanes, const double *in, double *out);

/* Bytes of the fixed device scratch of rd_sht_synthesis_noalloc (64 MiB). */
#define RD_NOALLOC_SCRATCH_BYTES (64ull << 20)

/* rd_sht_synthesis with no allocation of any kind: the Fourier scratch is a
 * fixed device array of RD_NOALLOC_SCRATCH_BYTES per GPU (held until the
 * process ends) instead of the plan's workspace; same kernels and results.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under stream capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The plan must exist (created beforehand).
 * Returns RD_ERR_ARG when 16 * nplanes * (lmax + 1) * nlat bytes exceed the
 * scratch.  All calls on one GPU share its scratch: calls that overlap on one
 * GPU (different streams, host threads) race. */
int rd_sht_synthesis_noalloc(rd_plan *plan, int nplanes, const double *in, double *out);

/* (v_r, v_theta, v_phi) -> (Q, S, T); S, T of l = 0 are 0.
 * Replaces fftp2d_vec_to_qst, src/sphp/sphp_vec.F90:49-106. */
int rd_vsht_analysis(rd_plan *plan, int nplanes, const double *vr, const double *vt,
                     const double *vp, double *q, double *s, double *t);

/* (Q, S, T) -> (v_r, v_theta, v_phi).
 * Replaces fftp2d_qst_to_vec, src/sphp/sphp_vec.F90:109-165. */
int rd_vsht_synthesis(rd_plan *plan, int nplanes, const double *q, const double *s,
                      const double *t, double *vr, double *vt, double *vp);

/* ---- radial (z) transforms, in place on (nlm, nz) ------------------- */

/* FC-Gram continuation of rows nz-C+1..nz, forward c2c FFT (sign -1) along
 * k, division by nz.  Replaces fftp1d_real_to_complex_z,
 * src/sphp/sphp_sca.F90:332-397. */
int rd_z_forward(rd_plan *plan, double *inout);

/* Backward c2c FFT (sign +1), unnormalized.  Replaces
 * fftp1d_complex_to_real_z, src/sphp/sphp_sca.F90:495-530. */
int rd_z_inverse(rd_plan *plan, double *inout);

/* 2D then z: in(nlat,nphi,nz) -> out(nlm,nz) in (lm,kz).
 * Replaces fftp3d_real_to_complex, src/sphp/sphp_sca.F90:237-270. */
int rd_3d_forward(rd_plan *plan, const double *in, double *out);

/* z (in place on in, which is destroyed) then 2D: in(nlm,nz) -> out.
 * Replaces fftp3d_complex_to_real, src/sphp/sphp_sca.F90:400-433. */
int rd_3d_inverse(rd_plan *plan, double *in, double *out);

/* The planes <-> lm transposes.  With the whole domain on one GPU the
 * reference's exchange reduces to copying the nplanes columns (its nprocs =
 * 1 path, src/sphp/sphp_sca.F90:634-635).  Replaces fftp3d_redist_k_to_lm /
 * fftp3d_redist_lm_to_k, src/sphp/sphp_sca.F90:629-760. */
int rd_redist_k_to_lm(rd_plan *plan, int nplanes, const double *in, double *out);
int rd_redist_lm_to_k(rd_plan *plan, int nplanes, const double *in, double *out);

/* ---- pseudospectral operators, src/pseudo/pseudospec_hd.f90 ---------- */

/* d^p a / dr^p; esp = 'r' (lm,r) or 'k' (lm,kz).  Replaces derivr /
 * derivr_r / derivr_k, pseudospec_hd.f90:18-135. */
int rd_derivr(rd_plan *plan, int p, char esp, const double *a, double *b);

/* b = -l(l+1) a.  Replaces laplang, pseudospec_hd.f90:138-172. */
int rd_laplang(rd_plan *plan, const double *a, double *b);

/* b = r^p a.  Replaces multr, pseudospec_hd.f90:175-208. */
int rd_multr(rd_plan *plan, int p, const double *a, double *b);

/* b = d2a/dr2 - l(l+1) a / r^2.  Replaces dlr, pseudospec_hd.f90:254-288. */
int rd_dlr(rd_plan *plan, const double *a, double *b);

/* z = a x + b y on (nlm, nz) complex.  Replaces saxpby_c, :291-326. */
int rd_saxpby_c(rd_plan *plan, double *z, const double *x, double a, const double *y, double b);

/* z = a x + b y on (nlat, nphi, nplanes) real.  Replaces saxpby_r, :329-366. */
int rd_saxpby_r(rd_plan *plan, int nplanes, double *z, const double *x, double a,
                const double *y, double b);

/* C = A x B pointwise on the physical planes (k < min(nplanes, nz - C)).
 * Replaces prodvec, pseudospec_hd.f90:211-251. */
int rd_prodvec(rd_plan *plan, int nplanes, const double *ar, const double *at, const double *ap,
               const double *br, const double *bt, const double *bp, double *cr, double *ct,
               double *cp);

/* Exponential radial filter in place; esp 'r' or 'k'.  Replaces fc_filter,
 * fc_filter_k, fc_filter_r, pseudospec_hd.f90:369-467. */
int rd_fc_filter(rd_plan *plan, char esp, double *a);

/* Wall value of every lm column of the spectral array a (lm, nz) (device)
 * from its normal derivative: row 0 (boun 5, the inner radius z(1)) or row
 * nz - C - 1 (boun 6, z(nz - C)) holds the ord-th outward normal derivative
 * on entry and the reconstructed value on exit, from the d - 1 matching rows
 * next to the wall and the weights neu of rd_load_neumann_tables (host, d
 * values).  Other boundaries or orders print the reference's message and
 * return RD_ERR_ARG.  Replaces neumann_reconstruct along z,
 * src/sphp/fcgram_mod.f90:374-517 (z is the radial row index here). */
int rd_neumann_reconstruct(rd_plan *plan, double *a, int boun, int ord, const double *neu);

/* As rd_neumann_reconstruct for the Robin condition f' + coef f = g with one
 * real coef per lm column (device array of nlm): the value is divided by
 * coef neu(d) + 1.  neu is the first-order table.  Replaces
 * robin_reconstruct along z, src/sphp/fcgram_mod.f90:520-650. */
int rd_robin_reconstruct(rd_plan *plan, double *a, int boun, const double *neu, const double *coef);

/* b = a on (nlm, nz).  Replaces copy, pseudospec_hd.f90:470-497. */
int rd_copy(rd_plan *plan, const double *a, double *b);

/* ---- the composed nonlinear term ------------------------------------ */

/* Builds everything precision form prec needs (the FP32 / BF16 tables, the
 * int8 digit tables and operand scratch of RD_PREC_OZAKI) now instead of on
 * the form's first call, so that later calls allocate nothing and can be
 * captured.  No counterpart in the reference (its plan is complete at
 * fftp3d_create_plan). */
int rd_plan_prepare(rd_plan *plan, int prec);

/* N = u x omega with u = curl curl (W r) + curl (Z r):  W, Z (nlm, nz) in
 * (lm, r) -> NQ, NS, NT (nlm, nz) on the physical rows (rows nr..nz-1 are
 * set to zero).  W and Z are not modified.  prec selects RD_PREC_*.
 * Composition of derivr, dlr, multr, laplang (pseudospec_hd.f90),
 * fftp2d_qst_to_vec and fftp2d_vec_to_qst (src/sphp/sphp_vec.F90:49-165) and
 * prodvec (pseudospec_hd.f90:211-251); the reference has no routine of its
 * own for it. */
int rd_nonlinear_uxw(rd_plan *plan, int prec, const double *W, const double *Z, double *NQ,
                     double *NS, double *NT);

/* Captures one rd_nonlinear_uxw(plan, prec, W, Z, NQ, NS, NT) into a hipGraph
 * held by the plan (replacing an earlier one); rd_plan_replay launches it on
 * the bound stream.  The graph reads and writes these arrays on every
 * replay.  No reference routine. */
int rd_nonlinear_uxw_capture(rd_plan *plan, int prec, const double *W, const double *Z, double *NQ, double *NS,
                             double *NT);
int rd_plan_replay(rd_plan *plan);

/* One stage of the explicit low-storage Runge-Kutta scheme of order ord
 * around the path (stage o = ord, ..., 1: a = a0 + dt/o rhs(a); the stage
 * the boundary step closes, vboundary.f90:59-69), for the incompressible
 * velocity u = curl curl (W r) + curl (Z r) with prognostic Z and
 * phi = D_l W:
 *   1. N = u x omega from (W, Z) (rd_nonlinear_uxw, form prec);
 *   2. Z   = Z0   + dt/o (r N_T + nu D_l Z),
 *      phi = phi0 + dt/o (d(r N_S)/dr - N_Q + nu D_l phi)
 *      (the radial and the radial-curl-curl parts of the curl of the
 *      momentum equation, pressure free);
 *   3. filter != 0: fc_filter(Z, 'r'), fc_filter(phi, 'r');
 *   4. v_imposebc_fast(Z, phi, W): no slip on both walls, W = M_l phi.
 * Z0, phi0: the values at the start of the step (may be Z, phi themselves
 * for the first stage).  Needs rd_bc_setup.  Uses the plan's spectral scratch.
 * No single reference routine: the composition of rd_nonlinear_uxw,
 * derivr/dlr/multr/saxpby_c (pseudospec_hd.f90), fc_filter
 * (pseudospec_hd.f90:369-467) and v_imposebc_fast (vboundary.f90:91-130). */
int rd_stage_uxw(rd_plan *plan, int prec, int o, double dt, double nu, int filter, const double *Z0,
                 const double *phi0, double *Z, double *phi, double *W);

/* One nonlinear step of the coupled state: velocity (poloidal W, toroidal Z),
 * magnetic field B = curl curl (Pb r) + curl (Tb r) and a temperature scalar
 * T, all (nlm, nz) in (lm, r).  Outputs (nlm, nz), rows nr..nz-1 zero:
 *   (NQ, NS, NT) = u x omega, (MQ, MS, MT) = u x B  (Q, S, T vectors),
 *   NH = u . grad T  (scalar; grad T has Q = dT/dr, S = T/r, T = 0).
 * Inputs are not modified.  prec selects RD_PREC_* for every Legendre
 * contraction.  Composition of the same reference operators as
 * rd_nonlinear_uxw plus fftp2d_real_to_complex_xy
 * (src/sphp/sphp_sca.F90:273-329) for the scalar; no reference routine. */
int rd_nonlinear_coupled(rd_plan *plan, int prec, const double *W, const double *Z, const double *Pb,
                         const double *Tb, const double *T, double *NQ, double *NS, double *NT,
                         double *MQ, double *MS, double *MT, double *NH);

/* ---- boundary conditions, src/boundary/ ----------------------------- */

/* a(lm,1) = a(lm,nz-C) = 0.  Replaces dirichlet, boundary_mod.f90:62-93. */
int rd_dirichlet(rd_plan *plan, double *a);

#define RD_BC_NOSLIP 0      /* 'noslip': W = dW/dr = Z = 0 (vboundary.f90:10-31) */

/* FC-Gram radial derivative matrix of order p on the N = nr physical points,
 * d[i + N*j] = d^p e_j/dr^p at r_i (device, N x N column major).  Replaces
 * derivr_mat, boundary_mod.f90:96-137. */
int rd_derivr_mat(rd_plan *plan, int p, double *d);

/* Wall types (RD_BC_*) and the per-l recovery matrices M_l = columns 3..N-2 of
 * the row-scaled B_l^-1, built and LU-factorised on the device in FP64.
 * Returns RD_ERR_LU (and prints the l) when a B_l is singular.  Replaces
 * setup_bc / v_setup / invDl_setup, boundary_mod.f90:35-58,140-241 and
 * vboundary.f90:10-56. */
int rd_bc_setup(rd_plan *plan, int bcint, int bcext);

/* Copies (stream-ordered, device to device) of the tables built by
 * rd_bc_setup: binv (N x (N-4) per l = 1..lmax, the module array binv of
 * boundary_mod.f90:19-22) and d2bor (4 x N, :27-28).  Either destination may
 * be NULL. */
int rd_bc_tables(rd_plan *plan, double *binv, double *d2bor);

/* W = M_l phi(3:N-2) for l >= 1; l = 0 and the continuation rows are zero.
 * Replaces invDl, boundary_mod.f90:244-292. */
int rd_invdl(rd_plan *plan, const double *phi, double *W);

/* Z = 0 at the walls, W = M_l phi, phi = D_l W.  Replaces v_imposebc,
 * vboundary.f90:59-88. */
int rd_v_imposebc(rd_plan *plan, double *tor, double *phi, double *pol);

/* As rd_v_imposebc, with phi recomputed only on rows 1, 2, N-1, N and the
 * l = 0 mode of phi set to zero.  Replaces v_imposebc_fast,
 * vboundary.f90:91-130. */
int rd_v_imposebc_fast(rd_plan *plan, double *tor, double *phi, double *pol);

/* ---- device selection and the host -> device registry ---------------
 * Replaces src/memmgt/gdevice_mod.f90 for the Fortran drop-in layer: a host
 * array gets a device copy registered by its host address; an interior
 * address (one plane of the array) resolves to the same offset. */

/* device = rank mod number of devices.  device_init, gdevice_mod.f90:112-141. */
int rd_device_init(int rank);
/* device_sync, gdevice_mod.f90:144-155. */
int rd_device_sync(void);
/* gdev_alloc, gdevice_mod.f90:158-193. */
int rd_dev_alloc(const void *host, size_t nbytes);
/* gdev_free, gdevice_mod.f90:196-217. */
int rd_dev_free(const void *host);
/* Device address of a registered host address (an interior address resolves
 * to the same offset in the device copy), NULL if none. */
void *rd_dev_lookup(const void *host);
/* 1-based position of the array registered at exactly this host address in
 * the registry, 0 if none (gdev_lookup, gdevice_mod.f90:220-235; positions
 * follow the host addresses and change as arrays come and go, as the
 * reference's registry slots do on gdev_free). */
int rd_dev_index(const void *host);
/* gdev_present, gdevice_mod.f90:305-317. */
int rd_dev_present(const void *host);
/* gdev_update_to / gdev_update_from, gdevice_mod.f90:238-279. */
int rd_dev_update_to(const void *host, size_t nbytes);
int rd_dev_update_from(void *host, size_t nbytes);
/* gdev_copy, gdevice_mod.f90:282-302. */
int rd_dev_copy(const void *hdst, const void *hsrc, size_t nbytes);

/* ---- batching form ------------------------------------------------------
 * Many independent shell problems (each its own lmax, nlat, nphi, nz, dz,
 * rint; one FC-Gram table C, d, dir for all) transformed together: the
 * reference's fftp3d_complex_to_real / fftp3d_real_to_complex
 * (src/sphp/sphp_sca.F90:237-270,400-433) applied to every problem, with the
 * Legendre stage of all problems in one persistent launch per direction.
 * prec is RD_PREC_FP64 or RD_PREC_FP32.  Arrays are device pointers, one per
 * problem: spectral (nz, nlm) complex, physical (nz, nphi, nlat) real. */
typedef struct rd_batch rd_batch;

int rd_batch_create(rd_batch **b, int nprob, const int *lmax, const int *nlat, const int *nphi, const int *nz,
                    int C, int d, const double *dz, const double *rint, const double *dir, int prec);
int rd_batch_destroy(rd_batch *b);
/* As rd_plan_set_stream: NULL selects the default stream. */
int rd_batch_set_stream(rd_batch *b, void *stream);
int rd_batch_sync(rd_batch *b);
/* info: nprob, prec, plane-tile blocks, workgroups per CU, CUs, synthesis
 * and analysis work items, captured (8 ints). */
int rd_batch_info(const rd_batch *b, int *info);
/* spec -> grid per problem; spec is not modified. */
int rd_batch_3d_inverse(rd_batch *b, const double *const *spec, double *const *grid);
/* grid -> spec per problem (FC-Gram continuation, forward radial FFT, 1/nz). */
int rd_batch_3d_forward(rd_batch *b, const double *const *grid, double *const *spec);
/* Captures the round trip spec -> grid -> spec_out of every problem into one
 * hipGraph (replayed by rd_batch_replay on the batch stream). */
int rd_batch_capture(rd_batch *b, const double *const *spec, double *const *grid, double *const *spec_out);
int rd_batch_replay(rd_batch *b);
/* Tiles of the persistent launches: plane-tile blocks (2 or 4) and workgroups
 * per CU (1..8).  Drops a captured graph. */
int rd_batch_config(rd_batch *b, int jb, int wg_per_cu);
/* Times the round trip (reps each) for every tile choice, keeps the fastest;
 * reports the fixed default (4, 2) and the chosen times and the choice. */
int rd_batch_tune(rd_batch *b, const double *const *spec, double *const *grid, double *const *spec_out,
                  int reps, double *ms_fixed, double *ms_tuned, int *choice);

/* The nonlinear-term form of the batch: rd_nonlinear_uxw of every problem,
 * with the radial step, the fused phi transforms and the cross product per
 * problem (spread over the lanes) and the Legendre stage of all problems
 * (u and omega syntheses, then the analysis) as one persistent launch per
 * direction.  prec RD_PREC_FP64 or RD_PREC_FP32.  Every problem needs nz > C.
 * The round-trip calls (rd_batch_3d_*, rd_batch_capture, rd_batch_tune)
 * return RD_ERR_STATE on such a batch; rd_batch_replay, rd_batch_config
 * (plane-tile blocks 1 or 2) and rd_batch_info apply.  No reference routine:
 * the composition of rd_nonlinear_uxw, applied to many problems. */
int rd_batch_create_uxw(rd_batch **b, int nprob, const int *lmax, const int *nlat, const int *nphi,
                        const int *nz, int C, int d, const double *dz, const double *rint, const double *dir,
                        int prec);
/* W, Z -> NQ, NS, NT per problem (arrays of nprob device pointers, each
 * (nlm, nz) complex); W and Z are not modified. */
int rd_batch_nonlinear_uxw(rd_batch *b, const double *const *W, const double *const *Z, double *const *NQ,
                           double *const *NS, double *const *NT);
/* Captures one evaluation of every problem into one hipGraph
 * (rd_batch_replay launches it on the bound stream). */
int rd_batch_capture_uxw(rd_batch *b, const double *const *W, const double *const *Z, double *const *NQ,
                         double *const *NS, double *const *NT);
/* As rd_batch_tune, timing evaluations; the fixed default is (1, 2). */
int rd_batch_tune_uxw(rd_batch *b, const double *const *W, const double *const *Z, double *const *NQ,
                      double *const *NS, double *const *NT, int reps, double *ms_fixed, double *ms_tuned,
                      int *choice);
/* Problems with lmax >= lmax_min run their own one-problem chain (the kernels
 * of rd_3d_inverse / rd_3d_forward at the batch's precision, or of
 * rd_nonlinear_uxw, over the batch's lanes) instead of sharing the persistent
 * launches; 0 (the default): all of them, INT_MAX: none.  Results are
 * bit-identical either way.  The tuners (rd_batch_tune, rd_batch_tune_uxw)
 * first time this threshold (the default, none, and each distinct lmax of the
 * batch) at the default tiles, whose time is their ms_fixed, then the tiles
 * for the threshold kept.  Drops a captured graph.  No reference routine: a
 * schedule of the batching form. */
int rd_batch_split(rd_batch *b, int lmax_min);
/* The threshold in force and the number of problems it sends to one-problem
 * chains. */
int rd_batch_get_split(const rd_batch *b, int *lmax_min, int *n_one);

/* Measured FP64 ceilings of the current device in TFLOP/s: a register-only
 * loop of v_mfma_f64_16x16x4_f64 (the instruction of the FP64 Legendre
 * kernels) and one of v_fma_f64, each over eight workgroups of four waves per
 * CU.  A diagnostic for the roofline of the FP64 Legendre kernels; no
 * reference routine. */
int rd_fp64_ceilings(double *mfma_tflops, double *valu_tflops);

/* The same for a register-only loop of v_mfma_f64_4x4x4_4b_f64 (four 4x4x4
 * blocks per instruction), which sustains more FP64 work on this device than
 * the 16x16x4 form but needs twice the operands per flop when it stands in
 * for a 16x16 tile.  A diagnostic; no reference routine. */
int rd_fp64_mfma4x4_ceiling(double *tflops);

#ifdef __cplusplus
}
#endif

#endif /* REDONDITO_H */
