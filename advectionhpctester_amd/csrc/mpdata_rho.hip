// Density step with the velocity predictor: the reference's
// mpdata3d_rho_gauge_gpubc (advection/src_algorithms/mpdata3d-rho-gauge-
// gpubc.f90:23-622) and its velprd_driver (velprd_driver.f90:14-295).
//
// The driver is host policy around the predictor kernels: it measures the
// Courant number (a device max-reduction read back to the host), picks the
// number of substeps or the adaptive time step, predicts the C-grid
// velocities of n+1/2 and rotates the time levels.  The density step is
// infinite-gauge MPDATA of x_in = x_incomp/h with those velocities and no
// rhr factor, which also returns the total mass fluxes (first-order plus
// limited antidiffusive) in u1..u3 and the derived fields x_incomp, pexr1,
// pexr2, rhoadv and rhr.  Open boundaries only, as the reference's
// single-process path.
//
// Kernels: the fused density step is the density mode of the tracer step's
// fixed-position k-march kernel (FusedDens in mpdata_fused.hip), which
// writes x and the mass fluxes out of place, followed by k_rho_copyback
// into the reference's in-place arrays (neighbouring tiles read x and u
// while tiles write).  For liner != 0, the context's per-stage path or
// arrays too large for 32-bit offsets the stage kernels run: the density
// upwind (k_rho_upwind: xant, the donor fluxes f*ns including the wall
// faces, rhoadv), the gauge pseudo-velocity, face and limiter stages shared
// with the tracer step, and the update (k_rho_update).  The predictor's
// last extrapolation runs fused with the level rotation (k_extrap_rotate).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "../../include/mpdata_hip.h"
#include "mpdata_ctx.hpp"
#include "mpdata_fused.hpp"
#include "mpdata_stages.hpp"
#include "mpdata_status.hpp"

using namespace mpdata;

namespace {

#define CELL_LOOP_HEAD                                       \
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x; \
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y; \
    const int k = 1 + blockIdx.z;                            \
    if (i > g.np || j > g.mp || k > g.lp) return;

#define S_(a, ii, jj, kk) a[off_raw<SCAL>(g, ii, jj, kk)]
#define X_(a, ii, jj, kk) a[off_raw<FACEX>(g, ii, jj, kk)]
#define Y_(a, ii, jj, kk) a[off_raw<FACEY>(g, ii, jj, kk)]
#define Z_(a, ii, jj, kk) a[off_raw<FACEZ>(g, ii, jj, kk)]

// Upwind pass of the density step (:110-208): rhoadv = x_incomp at level n,
// xant = x - (sum of donor-flux differences)/h, and the donor fluxes of the
// low faces (the last cell of a direction also writes the wall face; the
// sky wall flux -donor(x(lp-1),x(lp),u3(lp)) is the mirrored flux of the
// top cell).
template <typename T>
__global__ void __launch_bounds__(256)
    k_rho_upwind(Geo g, const T* __restrict__ u1, const T* __restrict__ u2,
                 const T* __restrict__ u3, const T* __restrict__ x, const T* __restrict__ h,
                 const T* __restrict__ xinc, T* __restrict__ xant, T* __restrict__ rhoadv,
                 T* __restrict__ f1ns, T* __restrict__ f2ns, T* __restrict__ f3ns) {
    CELL_LOOP_HEAD
    const bool gnd = k == 1, sky = k == g.lp;
    const T xc = S_(x, i, j, k);
    const T f1p = donor(xc, S_(x, i + 1, j, k), X_(u1, i + 1, j, k));
    const T f1 = donor(S_(x, i - 1, j, k), xc, X_(u1, i, j, k));
    const T f2p = donor(xc, S_(x, i, j + 1, k), Y_(u2, i, j + 1, k));
    const T f2 = donor(S_(x, i, j - 1, k), xc, Y_(u2, i, j, k));
    T f3p, f3;
    if (gnd) {
        f3p = donor(xc, S_(x, i, j, k + 1), Z_(u3, i, j, k + 1));
        f3 = -f3p;
    } else if (sky) {
        f3 = donor(S_(x, i, j, k - 1), xc, Z_(u3, i, j, k));
        f3p = -f3;
    } else {
        f3p = donor(xc, S_(x, i, j, k + 1), Z_(u3, i, j, k + 1));
        f3 = donor(S_(x, i, j, k - 1), xc, Z_(u3, i, j, k));
    }
    S_(xant, i, j, k) = (xc - (f1p - f1 + f2p - f2 + f3p - f3) / S_(h, i, j, k));
    S_(rhoadv, i, j, k) = S_(xinc, i, j, k);
    X_(f1ns, i, j, k) = f1;
    Y_(f2ns, i, j, k) = f2;
    Z_(f3ns, i, j, k) = f3;
    if (i == g.np) X_(f1ns, i + 1, j, k) = f1p;
    if (j == g.mp) Y_(f2ns, i, j + 1, k) = f2p;
    if (sky) Z_(f3ns, i, j, k + 1) = f3p;
}

// Update of the density step (:377-606).  LINER = 0: gauge limited fluxes
// (as the tracer step), x = xant - (sum of flux differences)/h, x_incomp =
// x*h, pexr1 = x_incomp*pexe, pexr2 = x_incomp, rhr = rhoadv/x_incomp and
// the total mass fluxes u = f*ns + F on the low faces; the wall faces are
// evaluated as the reference writes them, (f*ns + inflow part) - outflow
// part (:567-593).  LINER = 1: x = xant and u = f*ns (:595-606).
template <typename T, int LINER>
__global__ void __launch_bounds__(256)
    k_rho_update(Geo g, const T* __restrict__ v1, const T* __restrict__ v2,
                 const T* __restrict__ v3, const T* __restrict__ cp, const T* __restrict__ cn,
                 const T* __restrict__ xant, const T* __restrict__ h,
                 const T* __restrict__ rhoadv, const T* __restrict__ pexe,
                 const T* __restrict__ f1ns, const T* __restrict__ f2ns,
                 const T* __restrict__ f3ns, T* __restrict__ x, T* __restrict__ xinc,
                 T* __restrict__ rhr, T* __restrict__ pexr1, T* __restrict__ pexr2,
                 T* __restrict__ u1, T* __restrict__ u2, T* __restrict__ u3) {
    CELL_LOOP_HEAD
    const T hc = S_(h, i, j, k);
    if (LINER) {
        const T xn = S_(xant, i, j, k);
        const T xi = xn * hc;
        S_(x, i, j, k) = xn;
        S_(xinc, i, j, k) = xi;
        S_(pexr1, i, j, k) = xi * S_(pexe, i, j, k);
        S_(pexr2, i, j, k) = xi;
        S_(rhr, i, j, k) = S_(rhoadv, i, j, k) / xi;
        X_(u1, i, j, k) = X_(f1ns, i, j, k);
        Y_(u2, i, j, k) = Y_(f2ns, i, j, k);
        Z_(u3, i, j, k) = Z_(f3ns, i, j, k);
        if (i == g.np) X_(u1, i + 1, j, k) = X_(f1ns, i + 1, j, k);
        if (j == g.mp) Y_(u2, i, j + 1, k) = Y_(f2ns, i, j + 1, k);
        if (k == g.lp) Z_(u3, i, j, k + 1) = Z_(f3ns, i, j, k + 1);
        return;
    }
    const bool gnd = k == 1, sky = k == g.lp;
#define CP(ii, jj, kk) S_(cp, ii, jj, kk)
#define CN(ii, jj, kk) S_(cn, ii, jj, kk)
    const T F1 = gflux(X_(v1, i, j, k), CP(i, j, k), CN(i - 1, j, k), CP(i - 1, j, k), CN(i, j, k));
    const T F1p =
        gflux(X_(v1, i + 1, j, k), CP(i + 1, j, k), CN(i, j, k), CP(i, j, k), CN(i + 1, j, k));
    const T F2 = gflux(Y_(v2, i, j, k), CP(i, j, k), CN(i, j - 1, k), CP(i, j - 1, k), CN(i, j, k));
    const T F2p =
        gflux(Y_(v2, i, j + 1, k), CP(i, j + 1, k), CN(i, j, k), CP(i, j, k), CN(i, j + 1, k));
    T F3, F3p;
    if (gnd) {
        F3 = gflux(-Z_(v3, i, j, k + 1), CP(i, j, k), CN(i, j, k + 1), CP(i, j, k + 1),
                   CN(i, j, k));
        F3p = gflux(Z_(v3, i, j, k + 1), CP(i, j, k + 1), CN(i, j, k), CP(i, j, k),
                    CN(i, j, k + 1));
    } else if (sky) {
        F3 = gflux(Z_(v3, i, j, k), CP(i, j, k), CN(i, j, k - 1), CP(i, j, k - 1), CN(i, j, k));
        F3p = gflux(-Z_(v3, i, j, k), CP(i, j, k - 1), CN(i, j, k), CP(i, j, k), CN(i, j, k - 1));
    } else {
        F3 = gflux(Z_(v3, i, j, k), CP(i, j, k), CN(i, j, k - 1), CP(i, j, k - 1), CN(i, j, k));
        F3p = gflux(Z_(v3, i, j, k + 1), CP(i, j, k + 1), CN(i, j, k), CP(i, j, k),
                    CN(i, j, k + 1));
    }
    const T xn = (S_(xant, i, j, k) - (F1p - F1 + F2p - F2 + F3p - F3) / hc);
    const T xi = xn * hc;
    S_(x, i, j, k) = xn;
    S_(xinc, i, j, k) = xi;
    S_(pexr1, i, j, k) = xi * S_(pexe, i, j, k);
    S_(pexr2, i, j, k) = xi;
    S_(rhr, i, j, k) = S_(rhoadv, i, j, k) / xi;
    X_(u1, i, j, k) = X_(f1ns, i, j, k) + F1;
    Y_(u2, i, j, k) = Y_(f2ns, i, j, k) + F2;
    Z_(u3, i, j, k) = Z_(f3ns, i, j, k) + F3;
    if (i == g.np) {
        const T a = X_(v1, i + 1, j, k);
        X_(u1, i + 1, j, k) = X_(f1ns, i + 1, j, k) +
                              pp(a) * min3(T(1), CP(i + 1, j, k), CN(i, j, k)) -
                              pn(a) * min3(T(1), CP(i, j, k), CN(i + 1, j, k));
    }
    if (j == g.mp) {
        const T a = Y_(v2, i, j + 1, k);
        Y_(u2, i, j + 1, k) = Y_(f2ns, i, j + 1, k) +
                              pp(a) * min3(T(1), CP(i, j + 1, k), CN(i, j, k)) -
                              pn(a) * min3(T(1), CP(i, j, k), CN(i, j + 1, k));
    }
    if (sky) {
        const T a = -Z_(v3, i, j, k);
        Z_(u3, i, j, k + 1) = Z_(f3ns, i, j, k + 1) +
                              pp(a) * min3(T(1), CP(i, j, k - 1), CN(i, j, k)) -
                              pn(a) * min3(T(1), CP(i, j, k), CN(i, j, k - 1));
    }
#undef CP
#undef CN
}

// wadv = 0.5*wadv over k = 1-ih..lp+ih with the full i, j halo
// (velprd_driver.f90:250-252): the first (lp+2ih) planes of the z-face
// array, a contiguous prefix.
template <typename T>
__global__ void k_halve(T* __restrict__ a, long long n) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x)
        a[e] = T(0.5) * a[e];
}

bool open_all(const Geo& g) { return !g.px && !g.py && !g.pz; }

// The fused density step runs in the fixed-position kernel of the tracer
// step (FusedDens, mpdata_fused.hip) where its 32-bit offsets apply.
bool rho_fusable(const Geo& g) { return open_all(g) && FusedDens<double>::supports(g); }

// The fused step's results back into the reference's in-place arrays: x_in
// on the interior and the four wall lines (the cells the step defines), and
// the mass fluxes on faces 1..n+1 of their direction, interior range.
template <typename T>
__global__ void __launch_bounds__(256)
    k_rho_copyback(Geo g, const T* __restrict__ xn, const T* __restrict__ m1,
                   const T* __restrict__ m2, const T* __restrict__ m3, T* __restrict__ x,
                   T* __restrict__ u1, T* __restrict__ u2, T* __restrict__ u3) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // 0 .. np+1
    const int j = blockIdx.y * blockDim.y + threadIdx.y;  // 0 .. mp+1
    const int k = 1 + blockIdx.z;
    if (i > g.np + 1 || j > g.mp + 1) return;
    const bool in_i = i >= 1 && i <= g.np, in_j = j >= 1 && j <= g.mp;
    if ((in_i && in_j) || (in_j && !in_i) || (in_i && !in_j)) {
        const long long c = off_raw<SCAL>(g, i, j, k);
        x[c] = xn[c];
    }
    if (in_j && i >= 1) X_(u1, i, j, k) = X_(m1, i, j, k);
    if (in_i && j >= 1) Y_(u2, i, j, k) = Y_(m2, i, j, k);
    if (in_i && in_j) {
        Z_(u3, i, j, k) = Z_(m3, i, j, k);
        if (k == g.lp) Z_(u3, i, j, k + 1) = Z_(m3, i, j, k + 1);
    }
}

// velprd_traj0's extrapolation (module_velprd.f90:55-85) followed, per cell,
// by the time-level rotation (velprd_driver.f90:254-270): level n-1 of ox is
// read before the rotation overwrites it with ox(.,0)*h.  The twelve level
// stores are non-temporal: nothing reads them back before they leave L2.
template <typename T>
__global__ void __launch_bounds__(256)
    k_extrap_rotate(Geo g, T* __restrict__ ox, T* __restrict__ oy, T* __restrict__ oz,
                    T* __restrict__ u, T* __restrict__ v, T* __restrict__ w,
                    const T* __restrict__ h, T gc1, T gc2, T gc3, T r) {
    CELL_LOOP_HEAD
    const long long n = len_of<SCAL>(g);
    const long long c = off_raw<SCAL>(g, i, j, k);
    const T hh = h[c];
    {
        const T a0 = ox[c] * hh, a2 = ox[c + 2 * n];
        const T a1 = a0 + r * (a0 - a2);
        __builtin_nontemporal_store(T(0.5) * (a1 + a0) * gc1, &ox[c + n]);
        __builtin_nontemporal_store(ox[c] * hh, &ox[c + 2 * n]);
    }
    {
        const T a0 = oy[c] * hh, a2 = oy[c + 2 * n];
        const T a1 = a0 + r * (a0 - a2);
        __builtin_nontemporal_store(T(0.5) * (a1 + a0) * gc2, &oy[c + n]);
        __builtin_nontemporal_store(oy[c] * hh, &oy[c + 2 * n]);
    }
    {
        const T a0 = oz[c] * hh, a2 = oz[c + 2 * n];
        const T a1 = a0 + r * (a0 - a2);
        __builtin_nontemporal_store(T(0.5) * (a1 + a0) * gc3, &oz[c + n]);
        __builtin_nontemporal_store(oz[c] * hh, &oz[c + 2 * n]);
    }
    const T u0 = u[c], v0 = v[c], w0 = w[c];
    __builtin_nontemporal_store(u0 + r * (u0 - u[c + 2 * n]), &u[c + n]);
    __builtin_nontemporal_store(v0 + r * (v0 - v[c + 2 * n]), &v[c + n]);
    __builtin_nontemporal_store(w0 + r * (w0 - w[c + 2 * n]), &w[c + n]);
    __builtin_nontemporal_store(u0, &u[c + 2 * n]);
    __builtin_nontemporal_store(v0, &v[c + 2 * n]);
    __builtin_nontemporal_store(w0, &w[c + 2 * n]);
}

dim3 cell_block() { return dim3(64, 4, 1); }
dim3 cell_grid(const Geo& g) { return dim3((g.np + 63) / 64, (g.mp + 3) / 4, g.lp); }

// Argument checks of the density step, made before any GPU work (also
// before the predictor driver of the whole routine).
int rho_check(mpdata_ctx* c, int liner, const void* u1, const void* u2, const void* u3,
              const void* x, const void* xinc, const void* h, const void* rhoadv, const void* rhr,
              const void* pexe, const void* pexr1, const void* pexr2, const void* bcx,
              const void* bcy) {
    const Geo& g = c->g;
    if (!open_all(g))
        return fail(MPDATA_EUNSUPPORTED,
                    "density step: open boundaries only (the reference's single-process path)");
    if (any_null(u1, u2, u3, x, xinc, h, rhoadv, rhr, pexe, pexr1, pexr2, bcx, bcy))
        return fail(MPDATA_EINVAL, "density step: null field");
    const void* outs[] = {x, xinc, rhoadv, rhr, pexr1, pexr2};
    const void* all[] = {x, xinc, rhoadv, rhr, pexr1, pexr2, h, pexe};
    for (const void* o : outs) {
        int same = 0;
        for (const void* a : all) same += a == o;
        if (same != 1) return fail(MPDATA_EINVAL, "density step: scalar fields must be distinct");
    }
    const bool fused = !liner && c->path != MPDATA_PATH_STAGES && rho_fusable(g);
    if (c->path == MPDATA_PATH_FUSED && !fused)
        return fail(MPDATA_EUNSUPPORTED, "fused density step: liner = 0 and arrays below 2^31 "
                                         "elements only");
    return MPDATA_OK;
}

// The density step after the predictor (:104-622 without the driver call),
// arguments checked by rho_check.  ox(.,2) = ox(.,0)*h (:104-108) is not
// repeated here: the driver's level rotation has just written the same
// products.
template <typename T>
int rho_run(mpdata_ctx* c, int liner, T* u1, T* u2, T* u3, T* x, T* xinc, const T* h,
            T* rhoadv, T* rhr, const T* pexe, T* pexr1, T* pexr2, const T* bcx, const T* bcy,
            hipStream_t s) {
    using St = Stages<T>;
    const Geo& g = c->g;
    Scratch& S = c->prim;
    int rc;
    const bool fused = !liner && c->path != MPDATA_PATH_STAGES && rho_fusable(g);
    if (fused) {
        if ((rc = ensure(c, &S.xnext, SCAL, sizeof(T)))) return rc;
        for (int d = 0; d < 3; d++)
            if ((rc = ensure(c, &S.f[d], FACEX + d, sizeof(T)))) return rc;
        T *xn = (T*)S.xnext, *m1 = (T*)S.f[0], *m2 = (T*)S.f[1], *m3 = (T*)S.f[2];
        const DensIO<T> dio{xinc, pexe, rhoadv, rhr, pexr1, pexr2, m1, m2, m3};
        // the fast build under MPDATA_PATH_FAST, else the exactly rounded one
        if (c->path == MPDATA_PATH_FAST)
            HIPCHK(FusedDensFast<T>::step(g, x, xn, u1, u2, u3, h, bcx, bcy, dio, s));
        else
            HIPCHK(FusedDens<T>::step(g, x, xn, u1, u2, u3, h, bcx, bcy, dio, s));
        dim3 cgrid((g.np + 2 + 63) / 64, (g.mp + 2 + 3) / 4, g.lp);
        k_rho_copyback<T><<<cgrid, cell_block(), 0, s>>>(g, xn, m1, m2, m3, x, u1, u2, u3);
        HIPCHK(hipGetLastError());
        return MPDATA_OK;
    }
    if ((rc = ensure(c, &S.xant, SCAL, sizeof(T)))) return rc;
    for (int d = 0; d < 3; d++) {
        if ((rc = ensure(c, &S.fns[d], FACEX + d, sizeof(T)))) return rc;
        if (!liner && (rc = ensure(c, &S.v[d], FACEX + d, sizeof(T)))) return rc;
    }
    if (!liner) {
        if ((rc = ensure(c, &S.cp, SCAL, sizeof(T)))) return rc;
        if ((rc = ensure(c, &S.cn, SCAL, sizeof(T)))) return rc;
    }
    T* xant = (T*)S.xant;
    T *f1ns = (T*)S.fns[0], *f2ns = (T*)S.fns[1], *f3ns = (T*)S.fns[2];
    T *v1 = (T*)S.v[0], *v2 = (T*)S.v[1], *v3 = (T*)S.v[2];
    T *cp = (T*)S.cp, *cn = (T*)S.cn;
    HIPCHK(St::putbcinxh(g, x, bcx, bcy, s));                          // :113
    k_rho_upwind<T><<<cell_grid(g), cell_block(), 0, s>>>(g, u1, u2, u3, x, h, xinc, xant, rhoadv,
                                                          f1ns, f2ns, f3ns);
    HIPCHK(hipGetLastError());
    if (liner) {
        k_rho_update<T, 1><<<cell_grid(g), cell_block(), 0, s>>>(
            g, nullptr, nullptr, nullptr, nullptr, nullptr, xant, h, rhoadv, pexe, f1ns, f2ns, f3ns,
            x, xinc, rhr, pexr1, pexr2, u1, u2, u3);
        HIPCHK(hipGetLastError());
        return MPDATA_OK;
    }
    HIPCHK(St::halo_xy(g, xant, St::MIRROR, 1, s));                    // :216
    HIPCHK(St::pseudo_vel(g, MPDATA_GAUGE, u1, u2, u3, xant, h, v1, v2, v3, nullptr, nullptr,
                          nullptr, s));                                // :225-275
    HIPCHK(St::face_bc(g, v1, v2, v3, s));                             // :276-280
    HIPCHK(St::halo_xy(g, x, St::COPY, 0, s));                         // :286
    HIPCHK(St::limiter(g, x, xant, h, v1, v2, v3, cp, cn, s));         // :287-370
    HIPCHK(St::halo_xy(g, cp, St::ZERO, 1, s));                        // :375-376
    HIPCHK(St::halo_xy(g, cn, St::ZERO, 1, s));
    k_rho_update<T, 0><<<cell_grid(g), cell_block(), 0, s>>>(g, v1, v2, v3, cp, cn, xant, h,
                                                             rhoadv, pexe, f1ns, f2ns, f3ns, x,
                                                             xinc, rhr, pexr1, pexr2, u1, u2, u3);
    HIPCHK(hipGetLastError());
    return MPDATA_OK;
}

constexpr int kMaxTries = 1000;

template <typename T> struct Tiny;
template <> struct Tiny<double> { static constexpr double v = 2.2250738585072014e-308; };
template <> struct Tiny<float> { static constexpr float v = 1.17549435e-38f; };

template <typename T>
T maxr(T a, T b) { return a > b ? a : b; }
template <typename T>
T minr(T a, T b) { return a < b ? a : b; }

// Predictor driver: velprd_driver.f90:14-295.  ox..w are three time levels
// (n, n+1/2, n-1) of one scalar field each.  The scalars follow the
// reference's arithmetic in the field's real type, including its
// default-real literals 0.1 and 0.99 (:208, :226).
template <typename T>
struct Driver {
    mpdata_ctx* c;
    T *ox, *oy, *oz, *u1, *u2, *u3;
    const T* h;
    T dxi, dyi, dzi;
    hipStream_t s;
    T dt_ratio = T(1);
    T cr[7] = {};

    int traj0(T dtv) {  // module_velprd.f90:15-89
        const Geo& g = c->g;
        const long long n = len_of<SCAL>(g);
        HIPCHK(Stages<T>::velprd_extrap(g, ox, oy, oz, h, dtv * dxi, dtv * dyi, dtv * dzi, dt_ratio,
                                        s));
        HIPCHK(Stages<T>::a_to_c(g, ox + n, oy + n, oz + n, u1, u2, u3, s));
        return MPDATA_OK;
    }
    // The last prediction of the call fused with the level rotation that
    // follows it (the rotation then does not run again).
    bool rotated = false;
    int traj0_rotate(T dtv, T* u, T* v, T* w) {
        const Geo& g = c->g;
        const long long n = len_of<SCAL>(g);
        dim3 grid((g.np + 63) / 64, (g.mp + 3) / 4, g.lp);
        k_extrap_rotate<T><<<grid, dim3(64, 4, 1), 0, s>>>(g, ox, oy, oz, u, v, w, h, dtv * dxi,
                                                           dtv * dyi, dtv * dzi, dt_ratio);
        HIPCHK(hipGetLastError());
        HIPCHK(Stages<T>::a_to_c(g, ox + n, oy + n, oz + n, u1, u2, u3, s));
        rotated = true;
        return MPDATA_OK;
    }
    int a_to_c() {
        const long long n = len_of<SCAL>(c->g);
        HIPCHK(Stages<T>::a_to_c(c->g, ox + n, oy + n, oz + n, u1, u2, u3, s));
        return MPDATA_OK;
    }
    int courants(int level, T dtv) {
        double out[7];
        int rc = reduce_courants<T>(c, ox, oy, oz, h, level, dxi, dyi, dzi, dtv, out, s);
        for (int q = 0; q < 7; q++) cr[q] = (T)out[q];
        return rc;
    }
    T cr_for_dt(int lvertsplit) const {
        return lvertsplit ? maxr(cr[1], T(0.5) * cr[6]) + Tiny<T>::v : cr[0];
    }
};

template <typename T>
int driver_impl(mpdata_ctx* c, int* nsub, int* nsubo, int lsub, T* dtn, T* dtnold, int ladapt,
                int lvs, int lnph, T crmaxlim, T dtmax, T* ox, T* oy, T* oz, T* u, T* v, T* w,
                T* u1, T* u2, T* u3, const T* h, T dxi, T dyi, T dzi, T dt, hipStream_t s) {
    if (any_null(nsub, nsubo, dtn, dtnold, ox, oy, oz, u, v, w, u1, u2, u3, h))
        return fail(MPDATA_EINVAL, "predictor driver: null argument");
    if (lsub && ladapt)
        return fail(MPDATA_ESTOP, "substepping and adaptivity not supported together "
                                  "(velprd_driver.f90:67)");
    if (lsub && (*nsub < 1 || *nsubo < 1))
        return fail(MPDATA_EINVAL, "nsubsteps %d, nsubsteps_o %d must be >= 1", *nsub, *nsubo);
    Driver<T> d{c, ox, oy, oz, u1, u2, u3, h, dxi, dyi, dzi, s};
    int rc;
    T dtntemp = *dtn;
    if (lsub && *nsub == 1) d.dt_ratio = (dt / T(1)) / (dt / (T)*nsubo);  // :73
    if (lsub) dtntemp = dt;
    if (ladapt) dtntemp = *dtn;
    // :80-89.  With a fixed step the Courant number is only printed by the
    // reference, so it is not measured (no read-back, the call stays
    // asynchronous); the extrapolation into level 1 is kept, as its result
    // is the level the C-grid velocities come from.
    const bool need_cr = lsub || ladapt;
    if (lnph) {
        const Geo& g = c->g;
        HIPCHK(Stages<T>::velprd_extrap(g, ox, oy, oz, h, dtntemp * dxi, dtntemp * dyi,
                                        dtntemp * dzi, d.dt_ratio, s));
        if (need_cr && (rc = d.courants(2, dtntemp))) return rc;
    } else if (need_cr) {
        if ((rc = d.courants(1, dtntemp))) return rc;
    }
    T crdt = d.cr_for_dt(lvs);
    if (lsub) {                                                               // :100-186
        bool ok = true;
        if (crdt > crmaxlim && *nsub == 1) {
            *nsub = (int)std::ceil((double)crdt);
            if (crdt / (T)*nsub > crmaxlim) *nsub = *nsub + 1;
            ok = false;
        } else if (crdt / (T)*nsub > T(10) && *nsub > 1) {
            return fail(MPDATA_ESTOP, "Courant > 1 for nsubsteps > 1 (velprd_driver.f90:121-124)");
        }
        *dtn = dt / (T)*nsub;
        if (lvs) {                                                            // :135-142
            d.dt_ratio = *dtn / *dtnold;
            if ((rc = d.traj0(*dtn)) || (rc = d.courants(2, *dtn))) return rc;
            crdt = d.cr[0];
        }
        for (int tries = 0; !ok; tries++) {                                   // :143-166
            if (tries == kMaxTries)
                return fail(MPDATA_ESTOP, "no admissible substep count after %d trials", tries);
            d.dt_ratio = *dtn / *dtnold;
            if ((rc = d.traj0(*dtn)) || (rc = d.courants(2, *dtn))) return rc;
            crdt = d.cr_for_dt(lvs);
            if (crdt <= crmaxlim) {
                ok = true;
            } else {
                *nsub = *nsub + 1;
                *dtn = dt / (T)*nsub;
            }
        }
        *nsubo = *nsub;
        *dtnold = *dtn;
        *dtn = dt / (T)*nsub;
        if (*nsub > 1 || !lnph) rc = d.traj0_rotate(*dtn, u, v, w);
        else rc = d.a_to_c();
        if (rc) return rc;
    } else if (ladapt) {                                                      // :191-232
        T crtemp = crmaxlim;
        *dtnold = *dtn;
        for (int tries = 0;; tries++) {
            if (tries == kMaxTries)
                return fail(MPDATA_ESTOP, "no admissible time step after %d trials", tries);
            const T dtt = minr((crtemp / (std::fabs(crdt) + Tiny<T>::v)) * *dtn, dtmax);
            if (dtt > *dtnold && ((dtt - *dtnold) / *dtnold) > (T)0.1f) *dtn = dtt;
            if (dtt <= *dtnold) *dtn = dtt;
            d.dt_ratio = *dtn / *dtnold;
            if ((rc = d.traj0(*dtn)) || (rc = d.courants(2, *dtn))) return rc;
            crdt = d.cr_for_dt(lvs);
            if (crdt <= crmaxlim) break;
            crtemp = (T)0.99f * crtemp;
        }
    } else {                                                                  // :236-248
        rc = lnph ? d.a_to_c() : d.traj0_rotate(*dtn, u, v, w);
        if (rc) return rc;
    }
    const Geo& g = c->g;
    if (lvs) {                                                                // :250-252
        const long long n = nx_of<FACEZ>(g) * ny_of<FACEZ>(g) * (g.lp + 2LL * g.ih);
        const int blocks = (int)std::min<long long>(4096, (n + 255) / 256);
        k_halve<T><<<blocks, 256, 0, s>>>(u3, n);
        HIPCHK(hipGetLastError());
    }
    if (!d.rotated)
        HIPCHK(Stages<T>::velprd_rotate(g, u, v, w, ox, oy, oz, h, d.dt_ratio, s));  // :254-270
    return MPDATA_OK;
}

}  // namespace

// ============================================================================
extern "C" {

#define DEFINE_RHO_API(T, SFX, RB)                                                               \
    int mpdata_velprd_driver##SFX(mpdata_ctx* ctx, int* nsubsteps, int* nsubsteps_o,             \
                                  int lsubstepping, T* dtn, T* dtnold, int ltimeadapt,           \
                                  int lvertsplit, int lcrdiag_nphlf, T crmaxlim, T dtmax, T* ox, \
                                  T* oy, T* oz, T* u, T* v, T* w, T* uadv, T* vadv, T* wadv,     \
                                  const T* h, int ipoles0, int ibcx0, int ibcy0, int ibcz0,      \
                                  int np, int mp, int lp, int ih, T dxi, T dyi, T dzi, T dt,     \
                                  void* stream) {                                                \
        int rc;                                                                                  \
        if ((rc = check_ctx(ctx, RB))) return rc;                                                \
        if ((rc = check_dims(ctx, ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih))) return rc;     \
        return driver_impl<T>(ctx, nsubsteps, nsubsteps_o, lsubstepping, dtn, dtnold,            \
                              ltimeadapt, lvertsplit, lcrdiag_nphlf, crmaxlim, dtmax, ox, oy,    \
                              oz, u, v, w, uadv, vadv, wadv, h, dxi, dyi, dzi, dt,               \
                              (hipStream_t)stream);                                              \
    }                                                                                            \
    int mpdata_velprd_traj0_adapt_dt##SFX(mpdata_ctx* ctx, T* ox, T* oy, T* oz, const T* h,      \
                                          T dxi, T dyi, T dzi, T dt, T dt_ratio, void* stream) { \
        int rc;                                                                                  \
        if ((rc = check_ctx(ctx, RB))) return rc;                                                \
        if (any_null(ox, oy, oz, h)) return fail(MPDATA_EINVAL, "null field");                  \
        HIPCHK(Stages<T>::velprd_extrap(ctx->g, ox, oy, oz, h, dt * dxi, dt * dyi, dt * dzi,     \
                                        dt_ratio, (hipStream_t)stream));                         \
        return MPDATA_OK;                                                                        \
    }                                                                                            \
    int mpdata_rho_gauge_core##SFX(mpdata_ctx* ctx, int liner, T* u1, T* u2, T* u3, T* x_in,     \
                                   T* x_incomp, const T* h, T* rhoadv, T* rhr, const T* pexe,    \
                                   T* pexr1, T* pexr2, const T* bcx, const T* bcy,               \
                                   void* stream) {                                               \
        int rc;                                                                                  \
        if ((rc = check_ctx(ctx, RB))) return rc;                                                \
        if ((rc = rho_check(ctx, liner, u1, u2, u3, x_in, x_incomp, h, rhoadv, rhr, pexe, pexr1, \
                            pexr2, bcx, bcy)))                                                   \
            return rc;                                                                           \
        return rho_run<T>(ctx, liner, u1, u2, u3, x_in, x_incomp, h, rhoadv, rhr, pexe, pexr1,   \
                          pexr2, bcx, bcy, (hipStream_t)stream);                                 \
    }                                                                                            \
    int mpdata3d_rho_gauge_gpubc##SFX(                                                           \
        mpdata_ctx* ctx, int lupdatemulti, int* nsubsteps, int* nsubsteps_o, int lsubstepping,   \
        T* dtn, T* dtnold, int ltimeadapt, int lvertsplit, int lcrdiag_nphlf, T crmaxlim,        \
        T dtmax, T* ox, T* oy, T* oz, T* u, T* v, T* w, T* u1, T* u2, T* u3, T* x_in,            \
        T* x_incomp, const T* h, T* rhoadv, T* rhr, const T* pexe, T* pexr1, T* pexr2,           \
        const T* bcx, const T* bcy, int ipoles0, int ibcx0, int ibcy0, int ibcz0, int np,        \
        int mp, int lp, int ih, T dxi, T dyi, T dzi, T dt, int liner, void* stream) {            \
        (void)lupdatemulti;                                                                      \
        int rc;                                                                                  \
        if ((rc = check_ctx(ctx, RB))) return rc;                                                \
        if ((rc = check_dims(ctx, ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih))) return rc;     \
        if ((rc = rho_check(ctx, liner, u1, u2, u3, x_in, x_incomp, h, rhoadv, rhr, pexe, pexr1, \
                            pexr2, bcx, bcy)))                                                   \
            return rc;                                                                           \
        if ((rc = driver_impl<T>(ctx, nsubsteps, nsubsteps_o, lsubstepping, dtn, dtnold,         \
                                 ltimeadapt, lvertsplit, lcrdiag_nphlf, crmaxlim, dtmax, ox, oy, \
                                 oz, u, v, w, u1, u2, u3, h, dxi, dyi, dzi, dt,                  \
                                 (hipStream_t)stream)))                                          \
            return rc;                                                                           \
        return rho_run<T>(ctx, liner, u1, u2, u3, x_in, x_incomp, h, rhoadv, rhr, pexe, pexr1,   \
                          pexr2, bcx, bcy, (hipStream_t)stream);                                 \
    }

DEFINE_RHO_API(double, _f64, 8)
DEFINE_RHO_API(float, _f32, 4)

}  // extern "C"
