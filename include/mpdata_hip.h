/*
 * mpdata_hip.h -- C ABI of libmpdata_hip.so, the MI355X (gfx950) MPDATA
 * advection step.
 *
 * Drop-in boundary for the reference's advection driver call
 * (piotrows/advectionHPCtester, advection/src_drivers/advec.f90:116-244),
 * whose seam is Fortran procedure-pointer dispatch between
 * upwind3d_gpubc and antidiff3d_{gauge,standard}_gpubc (advec.f90:143-144).
 * A Fortran host binds these entry points through iso_c_binding
 * (advectionhpctester_amd/fortran/mpdata_hip.f90); Python binds them through
 * ctypes (advectionhpctester_amd/_lib.py).
 *
 * Conventions
 *  - Fields are the reference's Fortran arrays, column-major, lower bound
 *    1-ih in every direction (i fastest):
 *      scalar  x, xant, h(=rhoadv), rhr          (np+2ih)   (mp+2ih)   (lp+2ih)
 *      x-faces u1(=uadv)                         (np+1+2ih) (mp+2ih)   (lp+2ih)
 *      y-faces u2(=vadv)                         (np+2ih)   (mp+1+2ih) (lp+2ih)
 *      z-faces u3(=wadv)                         (np+2ih)   (mp+2ih)   (lp+1+2ih)
 *      bcx(mp,lp,2) = x_lr, bcy(np,lp,2) = x_bt  (no halo)
 *  - Unless stated otherwise every pointer is GPU memory (hipMalloc or
 *    managed) owned by the caller; the library owns only the scratch held
 *    by an mpdata_ctx.
 *  - 'stream' is a hipStream_t passed as void* (NULL = default stream).
 *    Calls are asynchronous on that stream unless marked "synchronous".
 *    A context is not re-entrant: calls sharing one context must be
 *    ordered (one stream, or explicit events).
 *  - Boundary flags ibcx/ibcy/ibcz: 0 = open (the reference's semantics),
 *    1 = periodic, defined here as a clean period-n wrap in each direction
 *    (the reference's single-process periodic path is incomplete).
 *  - Every function returns MPDATA_OK (0) or an error code; the message of
 *    the last error of the calling thread is mpdata_last_error().  Argument
 *    checks happen before any GPU work, so invalid calls have no effect.
 *  - _f64 functions need a context created with real_bytes = 8, _f32 with 4.
 */
#ifndef MPDATA_HIP_H
#define MPDATA_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define MPDATA_OK 0
#define MPDATA_EINVAL 1      /* invalid argument or shape mismatch        */
#define MPDATA_EHIP 2        /* HIP runtime error                         */
#define MPDATA_ENOMEM 3      /* device allocation failed                  */
#define MPDATA_EUNSUPPORTED 4 /* option the reference has no meaning for  */
#define MPDATA_ESTOP 5       /* a condition on which the reference STOPs  */

/* variant codes (advec.f90:148-211: legacyoptmode 6 -> UPWIND,
 * legacyoptmode 4 with t_adv 1 -> GAUGE, t_adv 2 -> STANDARD) */
#define MPDATA_UPWIND 0
#define MPDATA_GAUGE 1
#define MPDATA_STANDARD 2

/* execution paths of mpdata_step_* */
#define MPDATA_PATH_AUTO 0   /* fused kernel when it supports the options */
#define MPDATA_PATH_STAGES 1 /* one kernel per reference stage            */
#define MPDATA_PATH_FUSED 2  /* single-pass fused kernel (error if unsupported) */
/* The fused kernels' fast builds (FMA contraction, Newton-refined reciprocal
 * division; one each for the gauge, the standard and the density step) for
 * the single-domain step, step_oop, the time loop, CE_advection's fields
 * and the density step; not bit-identical to the reference (a few ulps per
 * step, of order 1e-14 normwise after 50 steps at 512x512x256, held to
 * 1e-12); error if unsupported.  The other entry points (stage calls, the
 * adjoint's recording) treat it as MPDATA_PATH_AUTO. */
#define MPDATA_PATH_FAST 3

typedef struct mpdata_ctx mpdata_ctx;

/* ---- library / context ------------------------------------------------ */

const char *mpdata_version(void);
/* Helpers of the Fortran drop-ins with array dummies: the address of an
 * array passed by reference, and whether a pointer is device or managed
 * memory (1) or host memory (0).  No counterparts in the reference. */
const void *mpdata_addr(const void *p);
int mpdata_is_device_ptr(const void *p);
const char *mpdata_last_error(void);

/* Creates the per-domain context (scratch for the antidiffusive passes) on
 * the current HIP device.  Replaces the module-global scratch of
 * common/scratch_datafields_modular.f90:282-334 and the single-process
 * set-up of geomset/init_subdomains (common/mpi_parallel.f90:88-501,
 * 606-682).  Requires np,mp,lp >= 3, ih >= 1, ibc* in {0,1},
 * real_bytes in {4,8}. */
int mpdata_ctx_create(int np, int mp, int lp, int ih, int ibcx, int ibcy, int ibcz,
                      int real_bytes, mpdata_ctx **out);
int mpdata_ctx_destroy(mpdata_ctx *ctx);
/* Selects the execution path of mpdata_step_* / mpdata_c_timeloop_*. */
int mpdata_ctx_set_path(mpdata_ctx *ctx, int path);
/* Bytes of device scratch currently held by the context. */
long long mpdata_ctx_scratch_bytes(const mpdata_ctx *ctx);

/* ---- stage-level drop-ins ------------------------------------------------
 * Same arguments, in the same order, as the reference subroutines, with the
 * context first and the stream last (the serialization flags are dropped).
 * np,mp,lp,ih and ibc* must match the context; ipoles0 must be 0 (the
 * Cartesian dwarf; the reference's pole code paths are commented out).  */

/* upwind3d_gpubc: advection/src_algorithms/upwind3d-gpubc.f90:20-28.
 * Writes the x/y halo of x_in from bcx/bcy on open walls (putbcinxh) and
 * the interior of xant.  xforc_impl/xforc_expl and dt are accepted and
 * unused, as in the reference (:93-100 commented out); they may be NULL. */
int mpdata_upwind3d_gpubc_f64(mpdata_ctx *ctx, const double *u1, const double *u2,
                              const double *u3, double *x_in, double *xant,
                              const double *xforc_impl, const double *xforc_expl,
                              const double *rhr, const double *h, const double *bcx,
                              const double *bcy, int ipoles0, int ibcx0, int ibcy0, int ibcz0,
                              int np, int mp, int lp, int ih, double dt, void *stream);
int mpdata_upwind3d_gpubc_f32(mpdata_ctx *ctx, const float *u1, const float *u2,
                              const float *u3, float *x_in, float *xant,
                              const float *xforc_impl, const float *xforc_expl,
                              const float *rhr, const float *h, const float *bcx,
                              const float *bcy, int ipoles0, int ibcx0, int ibcy0, int ibcz0,
                              int np, int mp, int lp, int ih, float dt, void *stream);

/* antidiff3d_gauge_gpubc: advection/src_algorithms/antidiff3d-gauge-gpubc.f90:14-16.
 * Corrective pass with the non-oscillatory limiter; updates x_in in place,
 * overwrites the x/y halos of xant (mirror) and x_in (edge copy) as the
 * reference does.  lupdatemulti and iflip are accepted and have no effect
 * in a single process, as in the reference. */
int mpdata_antidiff3d_gauge_gpubc_f64(mpdata_ctx *ctx, int lupdatemulti, const double *u1,
                                      const double *u2, const double *u3, double *xant,
                                      double *x_in, const double *rhr, const double *h,
                                      int iflip, int ipoles0, int ibcx0, int ibcy0, int ibcz0,
                                      int np, int mp, int lp, int ih, void *stream);
int mpdata_antidiff3d_gauge_gpubc_f32(mpdata_ctx *ctx, int lupdatemulti, const float *u1,
                                      const float *u2, const float *u3, float *xant,
                                      float *x_in, const float *rhr, const float *h, int iflip,
                                      int ipoles0, int ibcx0, int ibcy0, int ibcz0, int np,
                                      int mp, int lp, int ih, void *stream);

/* antidiff3d_standard_gpubc: advection/src_algorithms/antidiff3d-standard-gpubc.f90:14-16 */
int mpdata_antidiff3d_standard_gpubc_f64(mpdata_ctx *ctx, int lupdatemulti, const double *u1,
                                         const double *u2, const double *u3, double *xant,
                                         double *x_in, const double *rhr, const double *h,
                                         int iflip, int ipoles0, int ibcx0, int ibcy0,
                                         int ibcz0, int np, int mp, int lp, int ih,
                                         void *stream);
int mpdata_antidiff3d_standard_gpubc_f32(mpdata_ctx *ctx, int lupdatemulti, const float *u1,
                                         const float *u2, const float *u3, float *xant,
                                         float *x_in, const float *rhr, const float *h,
                                         int iflip, int ipoles0, int ibcx0, int ibcy0,
                                         int ibcz0, int np, int mp, int lp, int ih,
                                         void *stream);

/* ---- driver-level drop-in -------------------------------------------------
 * advec_dwarf: advection/src_drivers/advec.f90:116-136.  legacyoptmode 4
 * (two-step GPUBC; t_adv 1 gauge, 2 standard) or 6 (upwind only, then
 * x <- xant).  lvertsplit must be 0 (the split path is commented out in the
 * reference, :213-217).  x_lr/x_bt are the inflow values bcx/bcy.  The
 * optional trailing arrays of the reference (wadv2, rhr2, ...) are unused
 * there and are not part of this signature. */
int mpdata_advec_dwarf_f64(mpdata_ctx *ctx, int legacyoptmode, int t_adv, int lupdatemulti,
                           int lvertsplit, int np, int mp, int lp, int ih, int ipoles,
                           int ibcx, int ibcy, int ibcz, double *x, const double *x_lr,
                           const double *x_bt, const double *uadv, const double *vadv,
                           const double *wadv, const double *rhr, const double *rhoadv,
                           void *stream);
int mpdata_advec_dwarf_f32(mpdata_ctx *ctx, int legacyoptmode, int t_adv, int lupdatemulti,
                           int lvertsplit, int np, int mp, int lp, int ih, int ipoles,
                           int ibcx, int ibcy, int ibcz, float *x, const float *x_lr,
                           const float *x_bt, const float *uadv, const float *vadv,
                           const float *wadv, const float *rhr, const float *rhoadv,
                           void *stream);

/* ---- generalised step and time loops ---------------------------------------
 * One MPDATA time step: upwind followed by iters-1 corrective passes.
 * variant MPDATA_UPWIND ignores iters/nonos.  iters=2, nonos=1 is the
 * reference; iters>2 uses the limited velocity of the previous pass as the
 * advective velocity of the next and keeps x at time level n as limiter
 * bound; nonos=0 drops the limiter.  On return the interior of x holds the
 * new field; the halo of x is scratch. */
int mpdata_step_f64(mpdata_ctx *ctx, int variant, int iters, int nonos, double *x,
                    const double *u1, const double *u2, const double *u3, const double *rhr,
                    const double *h, const double *bcx, const double *bcy, void *stream);
int mpdata_step_f32(mpdata_ctx *ctx, int variant, int iters, int nonos, float *x,
                    const float *u1, const float *u2, const float *u3, const float *rhr,
                    const float *h, const float *bcx, const float *bcy, void *stream);

/* Out-of-place step: x_out <- MPDATA(x_in); x_in is left unchanged and
 * x_out receives the whole array as the in-place step would leave x (new
 * interior, halo as the reference leaves it).  x_in and x_out must not
 * overlap.  The fused kernel writes x_out directly, so a caller that
 * alternates two buffers moves each field through HBM once per step. */
int mpdata_step_oop_f64(mpdata_ctx *ctx, int variant, int iters, int nonos, const double *x_in,
                        double *x_out, const double *u1, const double *u2, const double *u3,
                        const double *rhr, const double *h, const double *bcx,
                        const double *bcy, void *stream);
int mpdata_step_oop_f32(mpdata_ctx *ctx, int variant, int iters, int nonos, const float *x_in,
                        float *x_out, const float *u1, const float *u2, const float *u3,
                        const float *rhr, const float *h, const float *bcx, const float *bcy,
                        void *stream);

/* nt steps of mpdata_step_* (the time loop of
 * advection_dwarf_test_modular.f90:185-223 without the norms).  On the
 * fused path the steps alternate between x and a context buffer, with one
 * final copy into x when nt is odd. */
int mpdata_c_timeloop_f64(mpdata_ctx *ctx, int variant, int iters, int nonos, int nt, double *x,
                          const double *u1, const double *u2, const double *u3,
                          const double *rhr, const double *h, const double *bcx,
                          const double *bcy, void *stream);
int mpdata_c_timeloop_f32(mpdata_ctx *ctx, int variant, int iters, int nonos, int nt, float *x,
                          const float *u1, const float *u2, const float *u3, const float *rhr,
                          const float *h, const float *bcx, const float *bcy, void *stream);

/* The reference's own C-ABI declaration (advection_dwarf_test_modular.f90:
 * 436-446, bind(c,name="upwind_c_timeloop"), never implemented there):
 * nt donor-cell upwind steps (each x <- upwind(x)) with zero open-boundary
 * inflow, FP64.  Accepts host or device pointers (the reference passes
 * C_LOC of host arrays); host data is staged through the GPU.  Synchronous. */
int upwind_c_timeloop(void *xtracer, void *uadv, void *vadv, void *wadv, void *rhr, void *h,
                      int np, int mp, int lp, int ih, int nt);

/* ---- velocity predictor: advection/src_algorithms/module_velprd.f90 -------
 * ox, oy, oz (and u, v, w) hold three consecutive scalar fields: time level
 * 0 (n), 1 (n+1/2 output) and 2 (n-1), i.e. the reference's (...,0:2). */

/* velprd_traj0 (:15-89): level 1 = extrapolation to n+1/2 in Courant units,
 * then the C-grid velocities u1,u2,u3 via velprdA_to_C. */
int mpdata_velprd_traj0_f64(mpdata_ctx *ctx, double *ox, double *oy, double *oz, double *u1,
                            double *u2, double *u3, const double *h, double dxi, double dyi,
                            double dzi, double dt, double dt_ratio, void *stream);
int mpdata_velprd_traj0_f32(mpdata_ctx *ctx, float *ox, float *oy, float *oz, float *u1,
                            float *u2, float *u3, const float *h, float dxi, float dyi,
                            float dzi, float dt, float dt_ratio, void *stream);
/* velprdA_to_C (:166-364) on single-level A-grid fields (their open-wall
 * halos are overwritten with the one-sided extrapolation). */
int mpdata_velprdA_to_C_f64(mpdata_ctx *ctx, double *ox, double *oy, double *oz, double *u1,
                            double *u2, double *u3, void *stream);
int mpdata_velprdA_to_C_f32(mpdata_ctx *ctx, float *ox, float *oy, float *oz, float *u1,
                            float *u2, float *u3, void *stream);
/* time-level rotation: advection/src_algorithms/velprd_driver.f90:254-270 */
int mpdata_velprd_rotate_f64(mpdata_ctx *ctx, double *u, double *v, double *w, double *ox,
                             double *oy, double *oz, const double *h, double dt_ratio,
                             void *stream);
int mpdata_velprd_rotate_f32(mpdata_ctx *ctx, float *u, float *v, float *w, float *ox, float *oy,
                             float *oz, const float *h, float dt_ratio, void *stream);

/* velprd_traj0 followed by velprdA_to_C (module_velprd.f90:15-89, 166-364)
 * in one pass, open walls, for a time loop that rotates the time levels by
 * swapping pointers: ox, oy, oz are level n (one 3-D array each) and oxp,
 * oyp, ozp the previous step's level-n arrays, which this call multiplies by
 * h to form level n-1 as the reference's rotation stores it (ox(.,2) =
 * ox(.,0)*h, velprd_driver.f90:254-270; h must be that rotation's h).
 * u1..u3 are bit-identical to velprd_traj0 + velprdA_to_C on the stored
 * levels; the A-grid level n+1/2 is not stored.  No counterpart of its own
 * in the reference (it fuses two of its subroutines and its rotation). */
int mpdata_velprd_c_f64(mpdata_ctx *ctx, const double *ox, const double *oy, const double *oz,
                        const double *oxp, const double *oyp, const double *ozp,
                        const double *h, double dxi, double dyi, double dzi, double dt,
                        double dt_ratio, double *u1, double *u2, double *u3, void *stream);
int mpdata_velprd_c_f32(mpdata_ctx *ctx, const float *ox, const float *oy, const float *oz,
                        const float *oxp, const float *oyp, const float *ozp, const float *h,
                        float dxi, float dyi, float dzi, float dt, float dt_ratio, float *u1,
                        float *u2, float *u3, void *stream);
/* velprd_traj0_adapt_dt (module_velprd.f90:91-164): the extrapolation of
 * velprd_traj0 into level 1 without the C-grid averaging. */
int mpdata_velprd_traj0_adapt_dt_f64(mpdata_ctx *ctx, double *ox, double *oy, double *oz,
                                     const double *h, double dxi, double dyi, double dzi,
                                     double dt, double dt_ratio, void *stream);
int mpdata_velprd_traj0_adapt_dt_f32(mpdata_ctx *ctx, float *ox, float *oy, float *oz,
                                     const float *h, float dxi, float dyi, float dzi, float dt,
                                     float dt_ratio, void *stream);

/* velprd_driver (advection/src_algorithms/velprd_driver.f90:14-295): the
 * Courant diagnosis, substepping (lsubstepping: nsubsteps and dtn chosen
 * so that the Courant number stays <= crmaxlim) or adaptive time step
 * (ltimeadapt: dtn <= dtmax), the predicted C-grid velocities uadv, vadv,
 * wadv of n+1/2 (halved w for lvertsplit) and the rotation of the time
 * levels of ox..w.  nsubsteps, nsubsteps_o, dtn, dtnold are host scalars,
 * read and updated as the reference's INOUT arguments.  With substepping
 * or adaptivity the Courant read-backs synchronise the stream; with a fixed
 * step (neither) the Courant number, which the reference only prints, is
 * not measured and the call is asynchronous.
 * MPDATA_ESTOP where the reference STOPs (substepping with adaptivity,
 * before any GPU work; Courant/nsubsteps > 10 while substepping) and where
 * its loops would not end (no admissible step within 1000 trials); the last
 * two are found after the Courant measurement, so level 1 of ox, oy, oz
 * and uadv..wadv may then hold a trial prediction. */
int mpdata_velprd_driver_f64(mpdata_ctx *ctx, int *nsubsteps, int *nsubsteps_o,
                             int lsubstepping, double *dtn, double *dtnold, int ltimeadapt,
                             int lvertsplit, int lcrdiag_nphlf, double crmaxlim, double dtmax,
                             double *ox, double *oy, double *oz, double *u, double *v, double *w,
                             double *uadv, double *vadv, double *wadv, const double *h,
                             int ipoles0, int ibcx0, int ibcy0, int ibcz0, int np, int mp,
                             int lp, int ih, double dxi, double dyi, double dzi, double dt,
                             void *stream);
int mpdata_velprd_driver_f32(mpdata_ctx *ctx, int *nsubsteps, int *nsubsteps_o,
                             int lsubstepping, float *dtn, float *dtnold, int ltimeadapt,
                             int lvertsplit, int lcrdiag_nphlf, float crmaxlim, float dtmax,
                             float *ox, float *oy, float *oz, float *u, float *v, float *w,
                             float *uadv, float *vadv, float *wadv, const float *h, int ipoles0,
                             int ibcx0, int ibcy0, int ibcz0, int np, int mp, int lp, int ih,
                             float dxi, float dyi, float dzi, float dt, void *stream);

/* ---- density step: advection/src_algorithms/mpdata3d-rho-gauge-gpubc.f90 --
 * mpdata3d_rho_gauge_gpubc (:23-622) with the reference's argument list
 * (serialization flags dropped, context first, stream last; liner by
 * value, as the reference never changes it): the predictor driver above,
 * then infinite-gauge MPDATA of x_in = x_incomp/h with the predicted u1..u3
 * and no rhr factor.  On return u1..u3 hold the total mass fluxes (donor
 * plus limited antidiffusive), x_in the new field, x_incomp = x_in*h,
 * pexr1 = x_incomp*pexe, pexr2 = x_incomp, rhoadv = x_incomp at level n and
 * rhr = rhoadv/x_incomp; liner != 0 stops after the upwind pass.  Open
 * boundaries only (MPDATA_EUNSUPPORTED otherwise).  x_in, x_incomp, rhoadv,
 * rhr, pexr1, pexr2, h and pexe must be distinct arrays. */
int mpdata3d_rho_gauge_gpubc_f64(
    mpdata_ctx *ctx, int lupdatemulti, int *nsubsteps, int *nsubsteps_o, int lsubstepping,
    double *dtn, double *dtnold, int ltimeadapt, int lvertsplit, int lcrdiag_nphlf,
    double crmaxlim, double dtmax, double *ox, double *oy, double *oz, double *u, double *v,
    double *w, double *u1, double *u2, double *u3, double *x_in, double *x_incomp,
    const double *h, double *rhoadv, double *rhr, const double *pexe, double *pexr1,
    double *pexr2, const double *bcx, const double *bcy, int ipoles0, int ibcx0, int ibcy0,
    int ibcz0, int np, int mp, int lp, int ih, double dxi, double dyi, double dzi, double dt,
    int liner, void *stream);
int mpdata3d_rho_gauge_gpubc_f32(
    mpdata_ctx *ctx, int lupdatemulti, int *nsubsteps, int *nsubsteps_o, int lsubstepping,
    float *dtn, float *dtnold, int ltimeadapt, int lvertsplit, int lcrdiag_nphlf,
    float crmaxlim, float dtmax, float *ox, float *oy, float *oz, float *u, float *v, float *w,
    float *u1, float *u2, float *u3, float *x_in, float *x_incomp, const float *h,
    float *rhoadv, float *rhr, const float *pexe, float *pexr1, float *pexr2, const float *bcx,
    const float *bcy, int ipoles0, int ibcx0, int ibcy0, int ibcz0, int np, int mp, int lp,
    int ih, float dxi, float dyi, float dzi, float dt, int liner, void *stream);
/* The density step after the predictor (mpdata3d-rho-gauge-gpubc.f90:104-
 * 622), with u1..u3 already holding the C-grid velocities: asynchronous. */
int mpdata_rho_gauge_core_f64(mpdata_ctx *ctx, int liner, double *u1, double *u2, double *u3,
                              double *x_in, double *x_incomp, const double *h, double *rhoadv,
                              double *rhr, const double *pexe, double *pexr1, double *pexr2,
                              const double *bcx, const double *bcy, void *stream);
int mpdata_rho_gauge_core_f32(mpdata_ctx *ctx, int liner, float *u1, float *u2, float *u3,
                              float *x_in, float *x_incomp, const float *h, float *rhoadv,
                              float *rhr, const float *pexe, float *pexr1, float *pexr2,
                              const float *bcx, const float *bcy, void *stream);

/* ---- diagnostics (synchronous) --------------------------------------------- */
/* compute_courants: common/diagutils.f90:588-672.  out[7] (host) = cr3d,
 * crxy, crxz, cryz, crx, cry, crz. */
int mpdata_compute_courants_f64(mpdata_ctx *ctx, const double *ox, const double *oy,
                                const double *oz, const double *h, int ntimelevel, double dxi,
                                double dyi, double dzi, double dt, double *out, void *stream);
int mpdata_compute_courants_f32(mpdata_ctx *ctx, const float *ox, const float *oy,
                                const float *oz, const float *h, int ntimelevel, float dxi,
                                float dyi, float dzi, float dt, double *out, void *stream);
/* No counterpart in the reference.  The Courant maxima of
 * mpdata_compute_courants_* for callers that cannot allocate device memory
 * (code under HIP graph capture, tight time loops, harnesses that forbid
 * allocation): no context, no hipMalloc/hipFree, asynchronous on `stream`.
 * The reduction partials live in a static device array of the library
 * (3.67 MB per GPU, held until the process ends), which bounds the domain:
 * ceil(np/64)*ceil(mp/4)*lp <= 65536 (e.g. 256x256x256); larger domains are
 * rejected with MPDATA_EINVAL.  out: DEVICE memory, 7 doubles (cr3d, crxy,
 * crxz, cryz, crx, cry, crz), the same values as mpdata_compute_courants_*.
 * Calls that overlap on one GPU share the array and race. */
int mpdata_compute_courants_noalloc_f64(const double *ox, const double *oy, const double *oz,
                                        const double *h, int ntimelevel, double dxi,
                                        double dyi, double dzi, double dt, int np, int mp,
                                        int lp, int ih, double *out, void *stream);
int mpdata_compute_courants_noalloc_f32(const float *ox, const float *oy, const float *oz,
                                        const float *h, int ntimelevel, float dxi, float dyi,
                                        float dzi, float dt, int np, int mp, int lp, int ih,
                                        double *out, void *stream);
/* ---- timers ------------------------------------------------------------------
 * Keyed by the reference's timer ids (common/mpi_parallel.f90:7088-7447:
 * ttini/ttbeg/ttend/ttprt), measured with HIP events on `stream`.  Off until
 * mpdata_ttini (which also resets them); then the library itself records
 * 55/57 (two-step gauge/standard driver, advec.f90:183-240) in
 * mpdata_advec_dwarf_*, and 16 (upwind) and 51/53 (gauge/standard
 * antidiffusive pass) around the per-stage passes; a host adds its own ids
 * (the dwarf's 300 around each advec_dwarf call).  mpdata_ttget returns an
 * id's call count and summed seconds; mpdata_ttprt prints the reference's
 * table (one process: average = maximum = minimum) and its sorted list.
 * Both synchronize the recorded events. */
int mpdata_ttini(void);
int mpdata_ttbeg(int id, void *stream);
int mpdata_ttend(int id, void *stream);
int mpdata_ttget(int id, int *calls, double *seconds);
int mpdata_ttprt(void);
/* ---- serialization savepoints --------------------------------------------
 * The analogue of the reference's serialbox calls
 * (advection_CEtimestep_twostep_GPUBC.f90:122-142, 355-375 under SERIALIZE):
 * open a serializer directory, then write named fields at named savepoints.
 * Each field becomes <dir>/<savepoint>__<name>.npy (NumPy format, Fortran
 * order, shape n1 x n2 x n3, no pickled objects) and one line of
 * <dir>/savepoints.txt ("<savepoint> <name> <file> <f8|f4> <n1> <n2> <n3>").
 * Fields may be device or host pointers; device fields are copied on
 * `stream`, which is synchronized.  Names: [A-Za-z0-9._-], 1-200 chars. */
int mpdata_serializer_open(const char *dir);
int mpdata_serializer_close(void);
int mpdata_serialize_field_f64(const char *savepoint, const char *name, const double *field,
                               int n1, int n2, int n3, void *stream);
int mpdata_serialize_field_f32(const char *savepoint, const char *name, const float *field,
                               int n1, int n2, int n3, void *stream);
/* Interior sums for the dwarf's norms (advection_test_setup.f90:472-492,
 * 595-617): out[4] (host) = sum x^2, sum x0^2, sum (x0-x)^2, sum x
 * (accumulated in FP64).  x0 may be NULL (its terms are then 0). */
int mpdata_norm_sums_f64(mpdata_ctx *ctx, const double *x, const double *x0, double *out,
                         void *stream);
int mpdata_norm_sums_f32(mpdata_ctx *ctx, const float *x, const float *x0, double *out,
                         void *stream);

/* ---- tangent-linear and adjoint steps --------------------------------------
 * No counterpart in the reference.  The tangent-linear step advances x and
 * its tangent dx (directional derivative along dx, du1..du3; NULL du = 0)
 * through mpdata_step's stage sequence on (value, tangent) pairs, for every
 * variant/iters/nonos the per-stage path supports.  Synchronous w.r.t. the
 * stream only. */
int mpdata_tl_step_f64(mpdata_ctx *ctx, int variant, int iters, int nonos, double *x, double *dx,
                       const double *u1, const double *u2, const double *u3, const double *du1,
                       const double *du2, const double *du3, const double *rhr, const double *h,
                       const double *bcx, const double *bcy, void *stream);
/* Adjoint of the non-oscillatory 2-iteration step, standard or gauge
 * variant (open walls, ih = 1, FP64 context).  mpdata_adj_reserve allocates
 * nslots tape slots; mpdata_adj_record_variant_f64 runs one forward step of
 * `variant` (MPDATA_STANDARD or MPDATA_GAUGE) x_in -> x_out keeping its
 * intermediates in `slot` (mpdata_adj_record_f64: the standard variant);
 * mpdata_adj_step_f64 maps xbar (adjoint of the interior of x_out) to the
 * adjoint of the interior of x_in, in place, and ADDS ufac times the
 * velocity adjoint to ubar1..ubar3 (zero them before the first step; ufac
 * lets a flow u(t) = f(t) u accumulate the gradient of its shape u).
 * x_out receives the new interior; its halo is not defined. */
int mpdata_adj_reserve(mpdata_ctx *ctx, int nslots);
int mpdata_adj_record_f64(mpdata_ctx *ctx, int slot, const double *x_in, double *x_out,
                          const double *u1, const double *u2, const double *u3,
                          const double *rhr, const double *h, const double *bcx,
                          const double *bcy, void *stream);
int mpdata_adj_record_variant_f64(mpdata_ctx *ctx, int slot, int variant, const double *x_in,
                                  double *x_out, const double *u1, const double *u2,
                                  const double *u3, const double *rhr, const double *h,
                                  const double *bcx, const double *bcy, void *stream);
int mpdata_adj_step_f64(mpdata_ctx *ctx, int slot, const double *u1, const double *u2,
                        const double *u3, const double *rhr, const double *h, double *xbar,
                        double *ubar1, double *ubar2, double *ubar3, double ufac, void *stream);

/* ---- many fields on one grid: CE_advection -----------------------------------
 * Replaces CE_advection (advection/src_drivers/advection_CEtimestep_twostep_GPUBC.f90:27-376):
 * nfields fields (the reference's th, u, v, w, tht, pstr and up to six moist
 * fields) advanced by one step with one velocity u1,u2,u3 (uadv, vadv,
 * wadv), one rhr and one h (rhoadv), on the context's grid.  Field f has its
 * inflow values bcx[f], bcy[f] (x_lr, x_bt) and role[f]: 1 for the fields
 * the reference passes to mpdatm3d_antidiff (u, v, w, tht, pstr), 0 for
 * those it passes to mpdata3d_antidiff (th, qv, qc, qr, qi, qs, qg).
 * alg_choice is mpdata_alg_choice: 1 all gauge, 2 all standard, 0 or > 2
 * role 1 gauge and role 0 standard (:145-155).  liner != 0: upwind only
 * (:330-341).  Each field's result equals mpdata_step_* (2 iterations,
 * limiter on, the variant above) on that field; the gauge fields run as one
 * launch of the batched fused kernel, reading u, h and rhr once per tile for
 * all of them.  lupdatemulti and iflip have no effect on one process
 * without poles and are not arguments.  The in-place form writes through a
 * context buffer per gauge field and copies back; the out-of-place form
 * writes x_out directly (no output may overlap an input or another output). */
#define MPDATA_MAX_FIELDS 64
int mpdata_ce_advection_f64(mpdata_ctx *ctx, int alg_choice, int liner, int nfields,
                            const int *role, double *const *x, const double *const *bcx,
                            const double *const *bcy, const double *u1, const double *u2,
                            const double *u3, const double *rhr, const double *h, void *stream);
int mpdata_ce_advection_f32(mpdata_ctx *ctx, int alg_choice, int liner, int nfields,
                            const int *role, float *const *x, const float *const *bcx,
                            const float *const *bcy, const float *u1, const float *u2,
                            const float *u3, const float *rhr, const float *h, void *stream);
int mpdata_ce_advection_oop_f64(mpdata_ctx *ctx, int alg_choice, int liner, int nfields,
                                const int *role, const double *const *x_in, double *const *x_out,
                                const double *const *bcx, const double *const *bcy,
                                const double *u1, const double *u2, const double *u3,
                                const double *rhr, const double *h, void *stream);
int mpdata_ce_advection_oop_f32(mpdata_ctx *ctx, int alg_choice, int liner, int nfields,
                                const int *role, const float *const *x_in, float *const *x_out,
                                const float *const *bcx, const float *const *bcy,
                                const float *u1, const float *u2, const float *u3,
                                const float *rhr, const float *h, void *stream);

/* ---- ensembles: many domains in one launch ---------------------------------
 * No counterpart in the reference (its batching is many fields on one grid,
 * advection_CEtimestep_twostep_GPUBC.f90).  n members of different shapes
 * (ih = 1, open walls), each advanced by the fused gauge step (2 iterations,
 * limiter on) in ONE persistent launch per step (plus one halo-copy launch);
 * a time loop can be captured into a HIP graph and launched once.  Results
 * are bit-identical to mpdata_step_oop_* on each member. */
typedef struct mpdata_ens mpdata_ens;
int mpdata_ens_create(int nmember, const int *np, const int *mp, const int *lp, int ih,
                      int real_bytes, mpdata_ens **out);
int mpdata_ens_destroy(mpdata_ens *ens);
/* Host arrays of nmember device pointers.  x alternates between xa (even
 * steps read it) and xb; the two must not overlap. */
int mpdata_ens_set_fields_f64(mpdata_ens *ens, double *const *xa, double *const *xb,
                              const double *const *u1, const double *const *u2,
                              const double *const *u3, const double *const *rhr,
                              const double *const *h, const double *const *bcx,
                              const double *const *bcy);
int mpdata_ens_set_fields_f32(mpdata_ens *ens, float *const *xa, float *const *xb,
                              const float *const *u1, const float *const *u2,
                              const float *const *u3, const float *const *rhr,
                              const float *const *h, const float *const *bcx,
                              const float *const *bcy);
/* Tile configurations of the batched kernel (0 .. mpdata_ens_nconfigs()-1)
 * and the k-chunk length (planes one work item marches). */
int mpdata_ens_nconfigs(void);
const char *mpdata_ens_config_name(int cfg);
int mpdata_ens_set_tiling(mpdata_ens *ens, int cfg, int clen);
int mpdata_ens_info(const mpdata_ens *ens, int *nitem, int *cfg, int *clen);
/* Which build of the fused step the batch runs: MPDATA_PATH_AUTO or
 * MPDATA_PATH_FUSED (the exactly rounded one, the default; results
 * bit-identical to mpdata_step_oop_*), or MPDATA_PATH_FAST (FMA contraction
 * and Newton-refined reciprocals; results bit-identical to mpdata_step_oop_*
 * on a context set to MPDATA_PATH_FAST, and within 1e-12 of the exact
 * build).  A recorded graph is dropped when the build changes. */
int mpdata_ens_set_path(mpdata_ens *ens, int path);
/* Times every (configuration, k-chunk length) on trial_steps steps of the
 * real fields and keeps the fastest; the choice is cached per set of
 * shapes for the process.  Reads xa and overwrites xb.  Synchronous. */
int mpdata_ens_autotune(mpdata_ens *ens, int trial_steps, void *stream, int *cfg, int *clen,
                        double *ms_per_step);
/* One step: parity 0 xa -> xb, 1 xb -> xa. */
int mpdata_ens_step(mpdata_ens *ens, int parity, void *stream);
/* nsteps steps starting from xa (the result is in xa if nsteps is even,
 * else in xb); use_graph = 1 records the loop once into a HIP graph (kept
 * until the fields, the tiling or nsteps change) and launches that graph. */
int mpdata_ens_run(mpdata_ens *ens, int nsteps, int use_graph, void *stream);

/* ---- device memory for hosts without a GPU allocator ------------------------
 * The reference declares its fields MANAGED and its scratch DEVICE through
 * CUDA Fortran attributes (advection/include/defines.inc:1-11).  A Fortran
 * host that links this library allocates its fields here instead and passes
 * the returned device addresses to the entry points above. */
int mpdata_dev_alloc(long long bytes, void **ptr);  /* zero-filled */
int mpdata_dev_free(void *ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device.  Synchronous. */
int mpdata_copy(void *dst, const void *src, long long bytes, int kind);
/* Waits for all work queued on 'stream' (NULL: the whole device). */
int mpdata_sync(void *stream);

#ifdef __cplusplus
}
#endif
#endif /* MPDATA_HIP_H */
