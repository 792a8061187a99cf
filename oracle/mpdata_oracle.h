/*
 * CPU oracle for the MPDATA advection step -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference algorithm (EULAG MPDATA
 * dwarf, piotrows/advectionHPCtester).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only as the checker.  The
 * product path (libmpdata_hip.so) never links, loads or calls anything here.
 *
 * Layout: every field is a Fortran column-major array with lower bound 1-ih
 * in each direction, exactly as the reference passes them:
 *   scalar  x,xant,h,rhr,cp,cn : (np+2ih) x (mp+2ih)   x (lp+2ih)
 *   x-faces u1,v1,f1,f1ns      : (np+1+2ih) x (mp+2ih) x (lp+2ih)
 *   y-faces u2,v2,f2,f2ns      : (np+2ih) x (mp+1+2ih) x (lp+2ih)
 *   z-faces u3,v3,f3,f3ns      : (np+2ih) x (mp+2ih)   x (lp+1+2ih)
 *   bcx(mp,lp,2), bcy(np,lp,2) : no halo
 *
 * Boundary flags ibcx/ibcy/ibcz: 0 = open (reference semantics), 1 =
 * periodic with a clean period-n wrap (build-defined; the reference's own
 * periodic path is incomplete in a single process).
 *
 * Every function has an _f64 (double) and an _f32 (float) instantiation.
 */
#ifndef MPDATA_ORACLE_H
#define MPDATA_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* Scratch arrays the antidiffusive passes use; all must be allocated by the
 * caller with the shapes listed above (f-arrays only for the standard
 * variant).  Exposed so tests can compare every stage, like the reference's
 * module-global scratch (scratch_datafields_modular.f90:282-334). */
typedef struct {
    void *v1, *v2, *v3;       /* antidiffusive pseudo-velocities (faces)  */
    void *cp, *cn;            /* limiter ratios beta-up / beta-down       */
    void *f1ns, *f2ns, *f3ns; /* standard: first-order antidiffusive flux */
    void *f1, *f2, *f3;       /* standard: limited donor fluxes; gauge:
                                 limited fluxes (written when non-NULL)   */
} orc_scratch;

#define ORC_DECL(R, S)                                                        \
void orc_putbcinxh##S(R *x, const R *bcx, const R *bcy, int ibcx, int ibcy,  \
                      int np, int mp, int lp, int ih);                       \
void orc_upwind3d_gpubc##S(const R *u1, const R *u2, const R *u3, R *x,      \
                      R *xant, const R *rhr, const R *h, const R *bcx,       \
                      const R *bcy, int ibcx, int ibcy, int ibcz,            \
                      int np, int mp, int lp, int ih);                       \
void orc_antidiff3d_gauge_gpubc##S(const R *u1, const R *u2, const R *u3,    \
                      R *xant, R *x, const R *rhr, const R *h,               \
                      int ibcx, int ibcy, int ibcz, int np, int mp, int lp,  \
                      int ih, orc_scratch *s);                               \
void orc_antidiff3d_standard_gpubc##S(const R *u1, const R *u2, const R *u3, \
                      R *xant, R *x, const R *rhr, const R *h,               \
                      int ibcx, int ibcy, int ibcz, int np, int mp, int lp,  \
                      int ih, orc_scratch *s);                               \
void orc_antidiff_pass##S(int variant, int nonos, const R *u1, const R *u2,  \
                      const R *u3, const R *psi, const R *x0, R *out,        \
                      const R *rhr, const R *h, int ibcx, int ibcy,          \
                      int ibcz, int np, int mp, int lp, int ih,              \
                      orc_scratch *s, R *ul1, R *ul2, R *ul3);               \
int  orc_mpdata_step##S(int variant, int iters, int nonos, R *x,             \
                      const R *u1, const R *u2, const R *u3, const R *rhr,   \
                      const R *h, const R *bcx, const R *bcy, int ibcx,      \
                      int ibcy, int ibcz, int np, int mp, int lp, int ih);   \
void orc_velprd_traj0##S(R *ox, R *oy, R *oz, R *u1, R *u2, R *u3,           \
                      const R *h, R dxi, R dyi, R dzi, R dt, R dt_ratio,     \
                      int ibcx, int ibcy, int ibcz, int np, int mp, int lp,  \
                      int ih);                                               \
void orc_velprdA_to_C##S(R *ox, R *oy, R *oz, R *u1, R *u2, R *u3,           \
                      int ibcx, int ibcy, int ibcz, int np, int mp, int lp,  \
                      int ih);                                               \
void orc_velprd_rotate##S(R *u, R *v, R *w, R *ox, R *oy, R *oz,             \
                      const R *h, R dt_ratio, int np, int mp, int lp,        \
                      int ih);                                               \
void orc_compute_courants##S(const R *ox, const R *oy, const R *oz,          \
                      const R *h, int ntimelevel, R dxi, R dyi, R dzi, R dt, \
                      int np, int mp, int lp, int ih, R out[7]);             \
void orc_velprd_traj0_adapt_dt##S(R *ox, R *oy, R *oz, const R *h, R dxi,    \
                      R dyi, R dzi, R dt, R dt_ratio, int np, int mp, int lp,\
                      int ih);                                               \
int  orc_velprd_driver##S(int *nsubsteps, int *nsubsteps_o,                  \
                      int lsubstepping, R *dtn, R *dtnold, int ltimeadapt,   \
                      int lvertsplit, int lcrdiag_nphlf, R crmaxlim,         \
                      R dtmax, R *ox, R *oy, R *oz, R *u, R *v, R *w, R *u1, \
                      R *u2, R *u3, const R *h, int ibcx, int ibcy, int ibcz,\
                      int np, int mp, int lp, int ih, R dxi, R dyi, R dzi,   \
                      R dt);                                                 \
int  orc_rho_gauge_core##S(int liner, R *u1, R *u2, R *u3, R *x_in,          \
                      R *x_incomp, const R *h, R *rhoadv, R *rhr,            \
                      const R *pexe, R *pexr1, R *pexr2, const R *bcx,       \
                      const R *bcy, int ibcx, int ibcy, int ibcz, int np,    \
                      int mp, int lp, int ih);                               \
int  orc_rho_gauge_step##S(int *nsubsteps, int *nsubsteps_o,                 \
                      int lsubstepping, R *dtn, R *dtnold, int ltimeadapt,   \
                      int lvertsplit, int lcrdiag_nphlf, R crmaxlim,         \
                      R dtmax, R *ox, R *oy, R *oz, R *u, R *v, R *w, R *u1, \
                      R *u2, R *u3, R *x_in, R *x_incomp, const R *h,        \
                      R *rhoadv, R *rhr, const R *pexe, R *pexr1, R *pexr2,  \
                      const R *bcx, const R *bcy, int ibcx, int ibcy,        \
                      int ibcz, int np, int mp, int lp, int ih, R dxi,       \
                      R dyi, R dzi, R dt, int liner);

ORC_DECL(double, _f64)
ORC_DECL(float, _f32)
#undef ORC_DECL

/* variant codes shared with the HIP library */
enum { ORC_UPWIND = 0, ORC_GAUGE = 1, ORC_STANDARD = 2 };

/* status codes of the predictor driver and the density step (0 = done,
 * -1 = allocation failure) */
enum { ORC_EPOLICY = 1, ORC_ECFL = 2, ORC_ENOCONV = 3, ORC_EPERIODIC = 4 };
#define ORC_MAXTRIES 1000

#ifdef __cplusplus
}
#endif
#endif
