/*
 * CPU oracle for the MPDATA step -- TEST INFRASTRUCTURE ONLY.
 * Instantiates mpdata_oracle_impl.inc for double and float.  Build without
 * fast-math and with -ffp-contract=off (oracle/Makefile) so every operation
 * is a single IEEE rounding, as in the reference's -O2 build.
 */
#include "mpdata_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>

typedef struct {
    int np, mp, lp, ih;
    int px, py, pz; /* periodic flags */
} orc_grid;

/* Linear offset of Fortran element (i,j,k), lower bound 1-ih, of an array of
 * the given staggering kind (0 scalar, 1/2/3 faces normal to x/y/z). */
static inline long orc_idx_raw(const orc_grid *g, int kind, int i, int j, int k)
{
    long nx = g->np + 2 * g->ih + (kind == 1);
    long ny = g->mp + 2 * g->ih + (kind == 2);
    return (long)(i + g->ih - 1) + nx * ((long)(j + g->ih - 1) + ny * (long)(k + g->ih - 1));
}

static inline int orc_wrap(int i, int n) { return ((i - 1) % n + n) % n + 1; }

/* As orc_idx_raw, with indices of periodic directions wrapped into 1..n
 * (cells) or 1..n (faces; face n+1 is face 1). */
static inline long orc_idx(const orc_grid *g, int kind, int i, int j, int k)
{
    if (g->px) i = orc_wrap(i, g->np);
    if (g->py) j = orc_wrap(j, g->mp);
    if (g->pz) k = orc_wrap(k, g->lp);
    return orc_idx_raw(g, kind, i, j, k);
}

static inline long orc_len(const orc_grid *g, int kind)
{
    return (long)(g->np + 2 * g->ih + (kind == 1)) * (g->mp + 2 * g->ih + (kind == 2)) *
           (g->lp + 2 * g->ih + (kind == 3));
}

static inline double fabsr_f64(double a) { return fabs(a); }
static inline double copysignr_f64(double a, double b) { return copysign(a, b); }
static inline float fabsr_f32(float a) { return fabsf(a); }
static inline float copysignr_f32(float a, float b) { return copysignf(a, b); }

/* ep_nonos = 1.e-10_euwp (common/epsilons.f90:4), HUGE() (diagutils.f90:605),
 * TINY() (velprd_driver.f90:90) */
#define R double
#define S _f64
#define ORC_EP 1.e-10
#define ORC_HUGE DBL_MAX
#define ORC_TINY DBL_MIN
#include "mpdata_oracle_impl.inc"
#undef R
#undef S
#undef ORC_EP
#undef ORC_HUGE
#undef ORC_TINY

#define R float
#define S _f32
#define ORC_EP 1.e-10f
#define ORC_HUGE FLT_MAX
#define ORC_TINY FLT_MIN
#include "mpdata_oracle_impl.inc"
#undef R
#undef S
#undef ORC_EP
#undef ORC_HUGE
#undef ORC_TINY
