// Per-stage MPDATA kernels for gfx950: one kernel per stage of the
// reference's two-step GPUBC path, used by the stage-level entry points
// (drop-ins for upwind3d_gpubc / antidiff3d_{gauge,standard}_gpubc) and as
// the building blocks of extended configurations (3+ iterations, unlimited
// passes, periodic boundaries).  The fused single-pass step lives in
// mpdata_fused.hip.
//
// Cell kernels run one thread per cell with 64x4 blocks (i fastest, one
// 64-wide wavefront per row segment), k on grid.z; the ground/sky plane
// special cases of the reference are branches on k that are uniform per
// block.
#include "mpdata_dual.hpp"
#include "mpdata_stages.hpp"

namespace mpdata {

#define CELL_LOOP_HEAD                                             \
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;       \
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;       \
    const int k = 1 + blockIdx.z;                                  \
    if (i > g.np || j > g.mp || k > g.lp) return;

#define S_(a, ii, jj, kk) a[off<SCAL>(g, ii, jj, kk)]
#define X_(a, ii, jj, kk) a[off<FACEX>(g, ii, jj, kk)]
#define Y_(a, ii, jj, kk) a[off<FACEY>(g, ii, jj, kk)]
#define Z_(a, ii, jj, kk) a[off<FACEZ>(g, ii, jj, kk)]

// ---- halo / boundary kernels (mpdataoperators.inc:253-287, 479-513,
//      551-584, 673-947, 1034-1063) --------------------------------------

// putbcinxh: x halo <- inflow values on open x/y walls.  2-D launch over the
// larger of the two wall extents; blockIdx.z selects the wall.
template <typename T>
__global__ void k_putbcinxh(Geo g, T* __restrict__ x, const T* __restrict__ bcx,
                            const T* __restrict__ bcy) {
    const int a = 1 + blockIdx.x * blockDim.x + threadIdx.x;  // j or i
    const int k = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    const int wall = blockIdx.z;                               // 0..3
    if (k > g.lp) return;
    if (wall < 2) {
        if (g.px || a > g.mp) return;
        x[off_raw<SCAL>(g, wall == 0 ? 0 : g.np + 1, a, k)] =
            bcx[(a - 1) + (long long)g.mp * ((k - 1) + (long long)g.lp * wall)];
    } else {
        if (g.py || a > g.np) return;
        x[off_raw<SCAL>(g, a, wall == 2 ? 0 : g.mp + 1, k)] =
            bcy[(a - 1) + (long long)g.np * ((k - 1) + (long long)g.lp * (wall - 2))];
    }
}

// Halo fills over the y walls (dir 1) or x walls (dir 0) with the full range
// of the other two directions when 'full', else the interior range.
//   mode 0: mirror    a(0)=a(2),   a(n+1)=a(n-1)   (cp_scnd_last_to_halo_xy_fulld)
//   mode 1: copy      a(0)=a(1),   a(n+1)=a(n)     (cp_last_to_halo_xy)
//   mode 2: zero      a(0)=0,      a(n+1)=0        (zero_halo_xyd)
template <typename T>
__global__ void k_halo_wall(Geo g, T* __restrict__ a, int dir, int mode, int full) {
    const int lo = full ? 1 - g.ih : 1;
    const int p = lo + blockIdx.x * blockDim.x + threadIdx.x;  // along the wall
    const int k = lo + blockIdx.y * blockDim.y + threadIdx.y;
    const int side = blockIdx.z;                                // 0 low, 1 high
    const int kmax = full ? g.lp + g.ih : g.lp;
    if (k > kmax) return;
    if (dir == 1) {  // y walls, p runs along i
        const int pmax = full ? g.np + g.ih : g.np;
        if (p > pmax) return;
        const int jd = side == 0 ? 0 : g.mp + 1;
        const int js = mode == 0 ? (side == 0 ? 2 : g.mp - 1) : (side == 0 ? 1 : g.mp);
        a[off_raw<SCAL>(g, p, jd, k)] = mode == 2 ? T(0) : a[off_raw<SCAL>(g, p, js, k)];
    } else {         // x walls, p runs along j
        const int pmax = full ? g.mp + g.ih : g.mp;
        if (p > pmax) return;
        const int id = side == 0 ? 0 : g.np + 1;
        const int is = mode == 0 ? (side == 0 ? 2 : g.np - 1) : (side == 0 ? 1 : g.np);
        a[off_raw<SCAL>(g, id, p, k)] = mode == 2 ? T(0) : a[off_raw<SCAL>(g, is, p, k)];
    }
}

// Open-wall conditions on C-grid arrays: x and y wall faces zeroed over the
// full halo range (flux_x_bcd / flux_y_bcd type 0; the y wall is row mp+1,
// where the reference writes row np+1, defines.inc:208-216), z walls
// mirrored a(1)=-a(2), a(lp+1)=-a(lp) (flux_z_bc_nohalod type 2).
template <typename T>
__global__ void k_face_bc_xy(Geo g, T* __restrict__ a1, T* __restrict__ a2) {
    const int p = 1 - g.ih + blockIdx.x * blockDim.x + threadIdx.x;
    const int k = 1 - g.ih + blockIdx.y * blockDim.y + threadIdx.y;
    const int side = blockIdx.z;  // 0,1: x walls of a1; 2,3: y walls of a2
    if (k > g.lp + g.ih) return;
    if (side < 2) {
        if (!a1 || g.px || p > g.mp + g.ih) return;
        a1[off_raw<FACEX>(g, side == 0 ? 1 : g.np + 1, p, k)] = T(0);
    } else {
        if (!a2 || g.py || p > g.np + g.ih) return;
        a2[off_raw<FACEY>(g, p, side == 2 ? 1 : g.mp + 1, k)] = T(0);
    }
}

template <typename T>
__global__ void k_face_bc_z(Geo g, T* __restrict__ a3) {
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    if (i > g.np || j > g.mp) return;
    a3[off_raw<FACEZ>(g, i, j, 1)] = -a3[off_raw<FACEZ>(g, i, j, 2)];
    a3[off_raw<FACEZ>(g, i, j, g.lp + 1)] = -a3[off_raw<FACEZ>(g, i, j, g.lp)];
}

// ---- donor-cell upwind: upwind3d-gpubc.f90:108-198 ------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_upwind(Geo g, const T* __restrict__ u1,
                                                const T* __restrict__ u2,
                                                const T* __restrict__ u3,
                                                const T* __restrict__ x, T* __restrict__ xant,
                                                const T* __restrict__ rhr,
                                                const T* __restrict__ h) {
    CELL_LOOP_HEAD
    const bool gnd = !g.pz && k == 1, sky = !g.pz && k == g.lp;
    T hinv = T(1) / S_(h, i, j, k);
    T xc = S_(x, i, j, k);
    T f1p = donor(xc, S_(x, i + 1, j, k), X_(u1, i + 1, j, k));
    T f1 = donor(S_(x, i - 1, j, k), xc, X_(u1, i, j, k));
    T f2p = donor(xc, S_(x, i, j + 1, k), Y_(u2, i, j + 1, k));
    T f2 = donor(S_(x, i, j - 1, k), xc, Y_(u2, i, j, k));
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
    S_(xant, i, j, k) = S_(rhr, i, j, k) * (xc - (f1p - f1 + f2p - f2 + f3p - f3) * hinv);
}

// ---- antidiffusive pseudo-velocities --------------------------------------
// gauge: antidiff3d-gauge-gpubc.f90:89-203; standard: antidiff3d-standard-
// gpubc.f90:72-202.  U = advective velocity of the pass.  Wall faces that the
// open boundary conditions overwrite are not evaluated.
template <typename T, int VARIANT>
__global__ void __launch_bounds__(256) k_pseudo_vel(Geo g, const T* __restrict__ u1,
                                                    const T* __restrict__ u2,
                                                    const T* __restrict__ u3,
                                                    const T* __restrict__ xa,
                                                    const T* __restrict__ h, T* __restrict__ v1,
                                                    T* __restrict__ v2, T* __restrict__ v3,
                                                    T* __restrict__ f1ns, T* __restrict__ f2ns,
                                                    T* __restrict__ f3ns) {
    CELL_LOOP_HEAD
    const bool gnd = !g.pz && k == 1, sky = !g.pz && k == g.lp;
    const bool zterms = !(gnd || sky);
    const T ep = Eps<T>::v;
#define XA(ii, jj, kk) S_(xa, ii, jj, kk)
#define AX(ii, jj, kk) fabs(S_(xa, ii, jj, kk))
    if (g.px || i > 1) {
        T hmx = T(1) / (T(0.5) * (S_(h, i - 1, j, k) + S_(h, i, j, k)));
        T s2 = Y_(u2, i - 1, j, k) + Y_(u2, i - 1, j + 1, k) + Y_(u2, i, j + 1, k) +
               Y_(u2, i, j, k);
        T val;
        if (VARIANT == 1) {
            T cr = s2 * rat4(XA(i - 1, j - 1, k), XA(i, j - 1, k), XA(i - 1, j + 1, k),
                             XA(i, j + 1, k));
            if (zterms) {
                T s3 = Z_(u3, i - 1, j, k) + Z_(u3, i - 1, j, k + 1) + Z_(u3, i, j, k + 1) +
                       Z_(u3, i, j, k);
                cr = cr + s3 * rat4(XA(i - 1, j, k - 1), XA(i, j, k - 1), XA(i - 1, j, k + 1),
                                    XA(i, j, k + 1));
            }
            val = vdyf(XA(i - 1, j, k), XA(i, j, k), X_(u1, i, j, k), hmx) -
                  T(0.125) * X_(u1, i, j, k) * hmx * (cr);
        } else {
            val = vdyfa(XA(i - 1, j, k), XA(i, j, k), X_(u1, i, j, k), hmx) +
                  vcorra(X_(u1, i, j, k), s2,
                         (AX(i - 1, j + 1, k) + AX(i, j + 1, k)) -
                             (AX(i - 1, j - 1, k) + AX(i, j - 1, k)),
                         AX(i - 1, j + 1, k) + AX(i, j + 1, k) + AX(i - 1, j - 1, k) +
                             AX(i, j - 1, k) + ep,
                         hmx);
            if (zterms) {
                T s3 = Z_(u3, i - 1, j, k) + Z_(u3, i - 1, j, k + 1) + Z_(u3, i, j, k + 1) +
                       Z_(u3, i, j, k);
                val = val + vcorra(X_(u1, i, j, k), s3,
                                   (AX(i - 1, j, k + 1) + AX(i, j, k + 1)) -
                                       (AX(i - 1, j, k - 1) + AX(i, j, k - 1)),
                                   AX(i - 1, j, k + 1) + AX(i, j, k + 1) + AX(i - 1, j, k - 1) +
                                       AX(i, j, k - 1) + ep,
                                   hmx);
            }
            X_(f1ns, i, j, k) = donor(XA(i - 1, j, k), XA(i, j, k), val);
        }
        X_(v1, i, j, k) = val;
    }
    if (g.py || j > 1) {
        T hmy = T(1) / (T(0.5) * (S_(h, i, j - 1, k) + S_(h, i, j, k)));
        T s1 = X_(u1, i, j - 1, k) + X_(u1, i, j, k) + X_(u1, i + 1, j, k) +
               X_(u1, i + 1, j - 1, k);
        T val;
        if (VARIANT == 1) {
            T cr = s1 * rat4(XA(i - 1, j - 1, k), XA(i - 1, j, k), XA(i + 1, j - 1, k),
                             XA(i + 1, j, k));
            if (zterms) {
                T s3 = Z_(u3, i, j - 1, k) + Z_(u3, i, j, k) + Z_(u3, i, j, k + 1) +
                       Z_(u3, i, j - 1, k + 1);
                cr = cr + s3 * rat4(XA(i, j - 1, k - 1), XA(i, j, k - 1), XA(i, j - 1, k + 1),
                                    XA(i, j, k + 1));
            }
            val = vdyf(XA(i, j - 1, k), XA(i, j, k), Y_(u2, i, j, k), hmy) -
                  T(0.125) * Y_(u2, i, j, k) * hmy * (cr);
        } else {
            val = vdyfa(XA(i, j - 1, k), XA(i, j, k), Y_(u2, i, j, k), hmy) +
                  vcorra(Y_(u2, i, j, k), s1,
                         (AX(i + 1, j - 1, k) + AX(i + 1, j, k)) -
                             (AX(i - 1, j - 1, k) + AX(i - 1, j, k)),
                         AX(i + 1, j - 1, k) + AX(i + 1, j, k) + AX(i - 1, j - 1, k) +
                             AX(i - 1, j, k) + ep,
                         hmy);
            if (zterms) {
                T s3 = Z_(u3, i, j - 1, k) + Z_(u3, i, j, k) + Z_(u3, i, j, k + 1) +
                       Z_(u3, i, j - 1, k + 1);
                val = val + vcorra(Y_(u2, i, j, k), s3,
                                   (AX(i, j - 1, k + 1) + AX(i, j, k + 1)) -
                                       (AX(i, j - 1, k - 1) + AX(i, j, k - 1)),
                                   AX(i, j - 1, k + 1) + AX(i, j, k + 1) + AX(i, j - 1, k - 1) +
                                       AX(i, j, k - 1) + ep,
                                   hmy);
            }
            Y_(f2ns, i, j, k) = donor(XA(i, j - 1, k), XA(i, j, k), val);
        }
        Y_(v2, i, j, k) = val;
    }
    if (!gnd) {
        T hmz = T(1) / (T(0.5) * (S_(h, i, j, k - 1) + S_(h, i, j, k)));
        T s1 = X_(u1, i, j, k - 1) + X_(u1, i, j, k) + X_(u1, i + 1, j, k) +
               X_(u1, i + 1, j, k - 1);
        T s2 = Y_(u2, i, j, k - 1) + Y_(u2, i, j + 1, k - 1) + Y_(u2, i, j + 1, k) +
               Y_(u2, i, j, k);
        T val;
        if (VARIANT == 1) {
            T cr = s1 * rat4(XA(i - 1, j, k - 1), XA(i - 1, j, k), XA(i + 1, j, k - 1),
                             XA(i + 1, j, k)) +
                   s2 * rat4(XA(i, j - 1, k - 1), XA(i, j - 1, k), XA(i, j + 1, k - 1),
                             XA(i, j + 1, k));
            val = vdyf(XA(i, j, k - 1), XA(i, j, k), Z_(u3, i, j, k), hmz) -
                  T(0.125) * Z_(u3, i, j, k) * hmz * (cr);
        } else {
            val = vdyfa(XA(i, j, k - 1), XA(i, j, k), Z_(u3, i, j, k), hmz) +
                  vcorra(Z_(u3, i, j, k), s1,
                         (AX(i + 1, j, k - 1) + AX(i + 1, j, k)) -
                             (AX(i - 1, j, k - 1) + AX(i - 1, j, k)),
                         AX(i + 1, j, k - 1) + AX(i + 1, j, k) + AX(i - 1, j, k - 1) +
                             AX(i - 1, j, k) + ep,
                         hmz) +
                  vcorra(Z_(u3, i, j, k), s2,
                         (AX(i, j + 1, k - 1) + AX(i, j + 1, k)) -
                             (AX(i, j - 1, k - 1) + AX(i, j - 1, k)),
                         AX(i, j + 1, k - 1) + AX(i, j + 1, k) + AX(i, j - 1, k - 1) +
                             AX(i, j - 1, k) + ep,
                         hmz);
            Z_(f3ns, i, j, k) = donor(XA(i, j, k - 1), XA(i, j, k), val);
        }
        Z_(v3, i, j, k) = val;
    }
#undef XA
#undef AX
}

// ---- non-oscillatory limiter ---------------------------------------------
// gauge antidiff3d-gauge-gpubc.f90:238-324, standard antidiff3d-standard-
// gpubc.f90:260-347 (w = f*ns).  On the ground plane the x bound is the
// 6-point minimum (the form of mpdata3d-rho-gauge-gpubc.f90:365-368); the
// gauge reference reads an unset mnijk_o there (:240-252).
template <typename T>
__global__ void __launch_bounds__(256) k_limiter(Geo g, const T* __restrict__ x0,
                                                 const T* __restrict__ xa,
                                                 const T* __restrict__ h,
                                                 const T* __restrict__ w1,
                                                 const T* __restrict__ w2,
                                                 const T* __restrict__ w3, T* __restrict__ cp,
                                                 T* __restrict__ cn) {
    CELL_LOOP_HEAD
    const bool gnd = !g.pz && k == 1, sky = !g.pz && k == g.lp;
    const T ep = Eps<T>::v;
    T mxo = S_(x0, i - 1, j, k), mno = mxo;
    T mx = S_(xa, i - 1, j, k), mn = mx;
#define ACC(a, lo, hi, ii, jj, kk)       \
    {                                    \
        T t_ = S_(a, ii, jj, kk);        \
        lo = min2(lo, t_);               \
        hi = max2(hi, t_);               \
    }
    ACC(x0, mno, mxo, i, j, k) ACC(x0, mno, mxo, i + 1, j, k)
    ACC(x0, mno, mxo, i, j - 1, k) ACC(x0, mno, mxo, i, j + 1, k)
    ACC(xa, mn, mx, i, j, k) ACC(xa, mn, mx, i + 1, j, k)
    ACC(xa, mn, mx, i, j - 1, k) ACC(xa, mn, mx, i, j + 1, k)
    if (!gnd) { ACC(x0, mno, mxo, i, j, k - 1) ACC(xa, mn, mx, i, j, k - 1) }
    if (!sky) { ACC(x0, mno, mxo, i, j, k + 1) ACC(xa, mn, mx, i, j, k + 1) }
#undef ACC
    mx = max2(mx, mxo);
    mn = min2(mn, mno);
    T a1p = X_(w1, i + 1, j, k), a1 = X_(w1, i, j, k);
    T a2p = Y_(w2, i, j + 1, k), a2 = Y_(w2, i, j, k);
    T a3p, a3;
    if (gnd) {
        a3p = Z_(w3, i, j, k + 1);
        a3 = -a3p;
    } else if (sky) {
        a3 = Z_(w3, i, j, k);
        a3p = -a3;
    } else {
        a3p = Z_(w3, i, j, k + 1);
        a3 = Z_(w3, i, j, k);
    }
    T xc = S_(xa, i, j, k), hc = S_(h, i, j, k);
    S_(cp, i, j, k) = (mx - xc) * hc / (pn(a1p) + pp(a1) + pn(a2p) + pp(a2) + pn(a3p) + pp(a3) + ep);
    S_(cn, i, j, k) = (xc - mn) * hc / (pp(a1p) + pn(a1) + pp(a2p) + pn(a2) + pp(a3p) + pn(a3) + ep);
}

// ---- gauge limited fluxes + update: antidiff3d-gauge-gpubc.f90:340-428 ----
// Without the limiter the flux is the pseudo-velocity.  When ul1 is given,
// the face fluxes (the limited velocities of the pass) are stored for a
// further pass; each cell writes its low face and the last cell of an open
// direction also the wall face.
template <typename T, int NONOS>
__global__ void __launch_bounds__(256) k_gauge_update(
    Geo g, const T* __restrict__ v1, const T* __restrict__ v2, const T* __restrict__ v3,
    const T* __restrict__ cp, const T* __restrict__ cn, const T* __restrict__ psi,
    const T* __restrict__ rhr, const T* __restrict__ h, T* __restrict__ out, T* __restrict__ ul1,
    T* __restrict__ ul2, T* __restrict__ ul3) {
    CELL_LOOP_HEAD
    const bool gnd = !g.pz && k == 1, sky = !g.pz && k == g.lp;
    T F1, F1p, F2, F2p, F3, F3p;
    if (NONOS) {
#define CP(ii, jj, kk) S_(cp, ii, jj, kk)
#define CN(ii, jj, kk) S_(cn, ii, jj, kk)
        F1 = gflux(X_(v1, i, j, k), CP(i, j, k), CN(i - 1, j, k), CP(i - 1, j, k), CN(i, j, k));
        F1p = gflux(X_(v1, i + 1, j, k), CP(i + 1, j, k), CN(i, j, k), CP(i, j, k),
                    CN(i + 1, j, k));
        F2 = gflux(Y_(v2, i, j, k), CP(i, j, k), CN(i, j - 1, k), CP(i, j - 1, k), CN(i, j, k));
        F2p = gflux(Y_(v2, i, j + 1, k), CP(i, j + 1, k), CN(i, j, k), CP(i, j, k),
                    CN(i, j + 1, k));
        if (gnd) {
            F3 = gflux(-Z_(v3, i, j, k + 1), CP(i, j, k), CN(i, j, k + 1), CP(i, j, k + 1),
                       CN(i, j, k));
            F3p = gflux(Z_(v3, i, j, k + 1), CP(i, j, k + 1), CN(i, j, k), CP(i, j, k),
                        CN(i, j, k + 1));
        } else if (sky) {
            F3 = gflux(Z_(v3, i, j, k), CP(i, j, k), CN(i, j, k - 1), CP(i, j, k - 1),
                       CN(i, j, k));
            F3p = gflux(-Z_(v3, i, j, k), CP(i, j, k - 1), CN(i, j, k), CP(i, j, k),
                        CN(i, j, k - 1));
        } else {
            F3 = gflux(Z_(v3, i, j, k), CP(i, j, k), CN(i, j, k - 1), CP(i, j, k - 1),
                       CN(i, j, k));
            F3p = gflux(Z_(v3, i, j, k + 1), CP(i, j, k + 1), CN(i, j, k), CP(i, j, k),
                        CN(i, j, k + 1));
        }
#undef CP
#undef CN
    } else {
        F1 = X_(v1, i, j, k);
        F1p = X_(v1, i + 1, j, k);
        F2 = Y_(v2, i, j, k);
        F2p = Y_(v2, i, j + 1, k);
        F3 = Z_(v3, i, j, k);
        F3p = Z_(v3, i, j, k + 1);
    }
    if (ul1) {
        X_(ul1, i, j, k) = F1;
        Y_(ul2, i, j, k) = F2;
        Z_(ul3, i, j, k) = F3;
        if (!g.px && i == g.np) X_(ul1, i + 1, j, k) = F1p;
        if (!g.py && j == g.mp) Y_(ul2, i, j + 1, k) = F2p;
        if (!g.pz && k == g.lp) Z_(ul3, i, j, k + 1) = F3p;
    }
    S_(out, i, j, k) = S_(rhr, i, j, k) * (S_(psi, i, j, k) / S_(rhr, i, j, k) -
                                           (F1p - F1 + F2p - F2 + F3p - F3) / S_(h, i, j, k));
}

// ---- standard: limited velocity + donor fluxes on interior faces ---------
// antidiff3d-standard-gpubc.f90:372-407 (ranges defines.inc:63-71,90-98,117-125)
template <typename T, int NONOS>
__global__ void __launch_bounds__(256) k_std_flux(
    Geo g, const T* __restrict__ v1, const T* __restrict__ v2, const T* __restrict__ v3,
    const T* __restrict__ cp, const T* __restrict__ cn, const T* __restrict__ psi,
    const T* __restrict__ rhr, T* __restrict__ f1, T* __restrict__ f2, T* __restrict__ f3,
    T* __restrict__ ul1, T* __restrict__ ul2, T* __restrict__ ul3) {
    CELL_LOOP_HEAD
    const T pc = S_(psi, i, j, k), rc = pc / S_(rhr, i, j, k);
    if (g.px || i > 1) {
        T v = X_(v1, i, j, k), vl = v;
        T pm = S_(psi, i - 1, j, k);
        if (NONOS) {
            T a1 = min3(T(1), S_(cp, i, j, k), S_(cn, i - 1, j, k));
            T a2 = min3(T(1), S_(cp, i - 1, j, k), S_(cn, i, j, k));
            vl = slimv(v, a1, a2, pm, pc);
        }
        if (ul1) X_(ul1, i, j, k) = vl;
        X_(f1, i, j, k) = donor(pm / S_(rhr, i - 1, j, k), rc, vl);
    }
    if (g.py || j > 1) {
        T v = Y_(v2, i, j, k), vl = v;
        T pm = S_(psi, i, j - 1, k);
        if (NONOS) {
            T a1 = min3(T(1), S_(cp, i, j, k), S_(cn, i, j - 1, k));
            T a2 = min3(T(1), S_(cp, i, j - 1, k), S_(cn, i, j, k));
            vl = slimv(v, a1, a2, pm, pc);
        }
        if (ul2) Y_(ul2, i, j, k) = vl;
        Y_(f2, i, j, k) = donor(pm / S_(rhr, i, j - 1, k), rc, vl);
    }
    if (g.pz || k > 1) {
        T v = Z_(v3, i, j, k), vl = v;
        T pm = S_(psi, i, j, k - 1);
        if (NONOS) {
            T a1 = min3(T(1), S_(cp, i, j, k), S_(cn, i, j, k - 1));
            T a2 = min3(T(1), S_(cp, i, j, k - 1), S_(cn, i, j, k));
            vl = slimv(v, a1, a2, pm, pc);
        }
        if (ul3) Z_(ul3, i, j, k) = vl;
        Z_(f3, i, j, k) = donor(pm / S_(rhr, i, j, k - 1), rc, vl);
    }
}

// ---- standard update: antidiff3d-standard-gpubc.f90:431-459 ---------------
template <typename T>
__global__ void __launch_bounds__(256) k_std_update(Geo g, const T* __restrict__ f1,
                                                    const T* __restrict__ f2,
                                                    const T* __restrict__ f3,
                                                    const T* __restrict__ psi,
                                                    const T* __restrict__ rhr,
                                                    const T* __restrict__ h,
                                                    T* __restrict__ out) {
    CELL_LOOP_HEAD
    const bool gnd = !g.pz && k == 1, sky = !g.pz && k == g.lp;
    T hinv = T(1) / S_(h, i, j, k);
    T sum = X_(f1, i + 1, j, k) - X_(f1, i, j, k) + Y_(f2, i, j + 1, k) - Y_(f2, i, j, k);
    if (gnd)
        sum = sum + Z_(f3, i, j, k + 1) + Z_(f3, i, j, k + 1);
    else if (sky)
        sum = sum - Z_(f3, i, j, k) - Z_(f3, i, j, k);
    else
        sum = sum + Z_(f3, i, j, k + 1) - Z_(f3, i, j, k);
    S_(out, i, j, k) = S_(rhr, i, j, k) * (S_(psi, i, j, k) / S_(rhr, i, j, k) - sum * hinv);
}

// x <- xant on the interior (advec.f90:225-232)
template <typename T>
__global__ void k_copy_interior(Geo g, const T* __restrict__ src, T* __restrict__ dst) {
    CELL_LOOP_HEAD
    S_(dst, i, j, k) = S_(src, i, j, k);
}

// ---- velocity predictor: module_velprd.f90 --------------------------------
// velprd_traj0 extrapolation (:60-83): level 1 = 0.5*(ox1+ox0)*gc
template <typename T>
__global__ void k_velprd_extrap(Geo g, T* __restrict__ ox, T* __restrict__ oy,
                                T* __restrict__ oz, const T* __restrict__ h, T gc1, T gc2, T gc3,
                                T dt_ratio) {
    CELL_LOOP_HEAD
    const long long n = len_of<SCAL>(g);
    const long long c = off_raw<SCAL>(g, i, j, k);
    T hh = h[c];
    T ox0 = ox[c] * hh, ox2 = ox[c + 2 * n];
    T ox1 = ox0 + dt_ratio * (ox0 - ox2);
    ox[c + n] = T(0.5) * (ox1 + ox0) * gc1;
    T oy0 = oy[c] * hh, oy2 = oy[c + 2 * n];
    T oy1 = oy0 + dt_ratio * (oy0 - oy2);
    oy[c + n] = T(0.5) * (oy1 + oy0) * gc2;
    T oz0 = oz[c] * hh, oz2 = oz[c + 2 * n];
    T oz1 = oz0 + dt_ratio * (oz0 - oz2);
    oz[c + n] = T(0.5) * (oz1 + oz0) * gc3;
}

// One-sided extrapolation into the open-wall halo (velprdA_to_C :222-264):
// a(0)=2a(1)-a(2), a(n+1)=2a(n)-a(n-1) along direction dir.
template <typename T>
__global__ void k_extrap_halo(Geo g, T* __restrict__ a, int dir) {
    const int p = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int q = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    const int side = blockIdx.z;
    if (dir == 0) {  // p=j, q=k
        if (p > g.mp || q > g.lp) return;
        const int n = g.np;
        if (side == 0)
            a[off_raw<SCAL>(g, 0, p, q)] = T(2) * a[off_raw<SCAL>(g, 1, p, q)] - a[off_raw<SCAL>(g, 2, p, q)];
        else
            a[off_raw<SCAL>(g, n + 1, p, q)] =
                T(2) * a[off_raw<SCAL>(g, n, p, q)] - a[off_raw<SCAL>(g, n - 1, p, q)];
    } else if (dir == 1) {  // p=i, q=k
        if (p > g.np || q > g.lp) return;
        const int n = g.mp;
        if (side == 0)
            a[off_raw<SCAL>(g, p, 0, q)] = T(2) * a[off_raw<SCAL>(g, p, 1, q)] - a[off_raw<SCAL>(g, p, 2, q)];
        else
            a[off_raw<SCAL>(g, p, n + 1, q)] =
                T(2) * a[off_raw<SCAL>(g, p, n, q)] - a[off_raw<SCAL>(g, p, n - 1, q)];
    } else {  // p=i, q=j
        if (p > g.np || q > g.mp) return;
        const int n = g.lp;
        if (side == 0)
            a[off_raw<SCAL>(g, p, q, 0)] = T(2) * a[off_raw<SCAL>(g, p, q, 1)] - a[off_raw<SCAL>(g, p, q, 2)];
        else
            a[off_raw<SCAL>(g, p, q, n + 1)] =
                T(2) * a[off_raw<SCAL>(g, p, q, n)] - a[off_raw<SCAL>(g, p, q, n - 1)];
    }
}

// A->C averaging u1(i)=0.5*(ox(i)+ox(i-1)) on faces 1..n+1 (open) or 1..n
// (periodic); launched over cells plus one extra layer.
template <typename T>
__global__ void k_a_to_c(Geo g, const T* __restrict__ ox, const T* __restrict__ oy,
                         const T* __restrict__ oz, T* __restrict__ u1, T* __restrict__ u2,
                         T* __restrict__ u3) {
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = 1 + blockIdx.z;
    const bool in_i = i <= g.np, in_j = j <= g.mp, in_k = k <= g.lp;
    if (in_j && in_k && i <= g.np + (g.px ? 0 : 1))
        X_(u1, i, j, k) = T(0.5) * (S_(ox, i, j, k) + S_(ox, i - 1, j, k));
    if (in_i && in_k && j <= g.mp + (g.py ? 0 : 1))
        Y_(u2, i, j, k) = T(0.5) * (S_(oy, i, j, k) + S_(oy, i, j - 1, k));
    if (in_i && in_j && k <= g.lp + (g.pz ? 0 : 1))
        Z_(u3, i, j, k) = T(0.5) * (S_(oz, i, j, k) + S_(oz, i, j, k - 1));
}

// The predictor in one pass (open walls): velprd_traj0's extrapolation of the
// A-grid velocity to n+1/2 (module_velprd.f90:15-89), the one-sided
// extrapolation into the wall halo and the averaging to the C grid
// (velprdA_to_C, :166-364), without storing the A-grid level n+1/2.  Level
// n-1 is read as the previous step's level-n array oxp times h, the value the
// reference's level rotation stores in ox(.,2) (velprd_driver.f90:254-270),
// so the caller rotates the time levels by swapping pointers; the products,
// sums and their order are those of k_velprd_extrap, k_extrap_halo and
// k_a_to_c, so u1..u3 are bit-identical to them as long as h is the same.
// One thread per (i, j, k) of the face ranges; the level-n+1/2 values of the
// neighbouring cells are recomputed from their inputs (cache hits).
template <typename T>
__device__ __forceinline__ T half_level(const T* __restrict__ a, const T* __restrict__ ap,
                                        const T* __restrict__ h, long long c, T gc, T r) {
    const T hh = h[c];
    const T a0 = a[c] * hh, a2 = ap[c] * hh;
    const T a1 = a0 + r * (a0 - a2);
    return T(0.5) * (a1 + a0) * gc;
}

// One thread per (i, j, k) of the face ranges; the level-n+1/2 values of the
// neighbouring cells are recomputed from their inputs (cache hits).  (A
// k-march with the z neighbour carried in registers measured slower: 226
// against 188 us at 256x256x128 in FP64, its serial plane loop exposing the
// load latency.)
template <typename T>
__global__ void __launch_bounds__(256)
k_predict_ac(Geo g, const T* __restrict__ ox, const T* __restrict__ oy, const T* __restrict__ oz,
             const T* __restrict__ oxp, const T* __restrict__ oyp, const T* __restrict__ ozp,
             const T* __restrict__ h, T gc1, T gc2, T gc3, T r, T* __restrict__ u1,
             T* __restrict__ u2, T* __restrict__ u3) {
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = 1 + blockIdx.z;
    const int np = g.np, mp = g.mp, lp = g.lp;
    auto cell = [&](int ii, int jj, int kk) { return off_raw<SCAL>(g, ii, jj, kk); };
    // level n+1/2 at cells m (upstream) and c (downstream) of a face at
    // position q along a direction of n cells: the wall faces use the
    // extrapolated halo cell 2a(1)-a(2) / 2a(n)-a(n-1)
    auto face = [&](const T* a, const T* ap, T gc, int q, int n, auto at) -> T {
        T lm, lc;
        if (q == 1) {
            const T l1 = half_level(a, ap, h, at(1), gc, r), l2 = half_level(a, ap, h, at(2), gc, r);
            lm = T(2) * l1 - l2;
            lc = l1;
        } else if (q == n + 1) {
            const T ln = half_level(a, ap, h, at(n), gc, r);
            const T lnm = half_level(a, ap, h, at(n - 1), gc, r);
            lm = ln;
            lc = T(2) * ln - lnm;
        } else {
            lm = half_level(a, ap, h, at(q - 1), gc, r);
            lc = half_level(a, ap, h, at(q), gc, r);
        }
        return T(0.5) * (lc + lm);
    };
    if (i <= np + 1 && j <= mp && k <= lp)
        u1[off_raw<FACEX>(g, i, j, k)] =
            face(ox, oxp, gc1, i, np, [&](int q) { return cell(q, j, k); });
    if (i <= np && j <= mp + 1 && k <= lp)
        u2[off_raw<FACEY>(g, i, j, k)] =
            face(oy, oyp, gc2, j, mp, [&](int q) { return cell(i, q, k); });
    if (i <= np && j <= mp && k <= lp + 1)
        u3[off_raw<FACEZ>(g, i, j, k)] =
            face(oz, ozp, gc3, k, lp, [&](int q) { return cell(i, j, q); });
}

// Time-level rotation: velprd_driver.f90:254-270
template <typename T>
__global__ void k_velprd_rotate(Geo g, T* __restrict__ u, T* __restrict__ v, T* __restrict__ w,
                                T* __restrict__ ox, T* __restrict__ oy, T* __restrict__ oz,
                                const T* __restrict__ h, T r) {
    CELL_LOOP_HEAD
    const long long n = len_of<SCAL>(g);
    const long long c = off_raw<SCAL>(g, i, j, k);
    u[c + n] = u[c] + r * (u[c] - u[c + 2 * n]);
    v[c + n] = v[c] + r * (v[c] - v[c + 2 * n]);
    w[c + n] = w[c] + r * (w[c] - w[c + 2 * n]);
    u[c + 2 * n] = u[c];
    v[c + 2 * n] = v[c];
    w[c + 2 * n] = w[c];
    ox[c + 2 * n] = ox[c] * h[c];
    oy[c + 2 * n] = oy[c] * h[c];
    oz[c + 2 * n] = oz[c] * h[c];
}

// ---- reductions ------------------------------------------------------------
// Courant maxima (diagutils.f90:588-672) and solution norms
// (advection_test_setup.f90:472-492, 595-617).  Each block reduces in LDS and
// writes one partial; the host finishes the tiny second level.
template <typename T>
__global__ void __launch_bounds__(256) k_courants(Geo g, const T* __restrict__ ox,
                                                  const T* __restrict__ oy,
                                                  const T* __restrict__ oz,
                                                  const T* __restrict__ h, int level, T gc1,
                                                  T gc2, T gc3, T* __restrict__ part) {
    __shared__ T red[7][256];
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = 1 + blockIdx.z;
    T m[7];
    for (int q = 0; q < 7; q++) m[q] = T(0);
    if (i <= g.np && j <= g.mp && k <= g.lp) {
        const long long n = len_of<SCAL>(g);
        const long long c = off_raw<SCAL>(g, i, j, k);
        T a, b, cc;
        if (level == 1) {
            a = fabs(gc1 * ox[c]);
            b = fabs(gc2 * oy[c]);
            cc = fabs(gc3 * oz[c]);
        } else {
            T dinv = T(1) / h[c];
            a = fabs(ox[c + n] * dinv);
            b = fabs(oy[c + n] * dinv);
            cc = fabs(oz[c + n] * dinv);
        }
        m[0] = a + b + cc; m[1] = a + b; m[2] = a + cc; m[3] = b + cc;
        m[4] = a; m[5] = b; m[6] = cc;
    }
    for (int q = 0; q < 7; q++) red[q][t] = m[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s)
            for (int q = 0; q < 7; q++) red[q][t] = max2(red[q][t], red[q][t + s]);
        __syncthreads();
    }
    if (t == 0) {
        const long long b = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
        for (int q = 0; q < 7; q++) part[7 * b + q] = red[q][0];
    }
}

// Second stage of the Courant maxima: the nb per-block partials (7 each)
// reduced by one workgroup into out[7].  A maximum does not depend on the
// order of its operands, so the result equals a host loop over the partials.
template <typename T>
__global__ void __launch_bounds__(256) k_max7(const T* __restrict__ part, long long nb,
                                              T* __restrict__ out) {
    __shared__ T red[7][256];
    const int t = threadIdx.x;
    T m[7];
    for (int q = 0; q < 7; q++) m[q] = T(0);
    for (long long b = t; b < nb; b += 256)
        for (int q = 0; q < 7; q++) m[q] = max2(m[q], part[7 * b + q]);
    for (int q = 0; q < 7; q++) red[q][t] = m[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s)
            for (int q = 0; q < 7; q++) red[q][t] = max2(red[q][t], red[q][t + s]);
        __syncthreads();
    }
    if (t < 7) out[t] = red[t][0];
}

// Sums of x^2, x0^2, (x0-x)^2 and x over the interior, in double.
template <typename T>
__global__ void __launch_bounds__(256) k_norm_sums(Geo g, const T* __restrict__ x,
                                                   const T* __restrict__ x0,
                                                   double* __restrict__ part) {
    __shared__ double red[4][256];
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = 1 + blockIdx.z;
    double s[4] = {0, 0, 0, 0};
    if (i <= g.np && j <= g.mp && k <= g.lp) {
        const long long c = off_raw<SCAL>(g, i, j, k);
        double a = (double)x[c], b = x0 ? (double)x0[c] : 0.0;
        s[0] = a * a; s[1] = b * b; s[2] = (b - a) * (b - a); s[3] = a;
    }
    for (int q = 0; q < 4; q++) red[q][t] = s[q];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (t < st)
            for (int q = 0; q < 4; q++) red[q][t] += red[q][t + st];
        __syncthreads();
    }
    if (t == 0) {
        const long long b = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
        for (int q = 0; q < 4; q++) part[4 * b + q] = red[q][0];
    }
}

#undef S_
#undef X_
#undef Y_
#undef Z_
#undef CELL_LOOP_HEAD

// ---- host-side launchers ----------------------------------------------------
namespace {
inline dim3 cell_block() { return dim3(64, 4, 1); }
inline dim3 cell_grid(const Geo& g, int extra = 0) {
    return dim3((g.np + extra + 63) / 64, (g.mp + extra + 3) / 4, g.lp + extra);
}
inline dim3 wall_block() { return dim3(64, 4, 1); }
}  // namespace

template <typename T>
hipError_t Stages<T>::putbcinxh(const Geo& g, T* x, const T* bcx, const T* bcy, hipStream_t s) {
    if (g.px && g.py) return hipSuccess;
    int a = g.mp > g.np ? g.mp : g.np;
    dim3 grid((a + 63) / 64, (g.lp + 3) / 4, 4);
    k_putbcinxh<T><<<grid, wall_block(), 0, s>>>(g, x, bcx, bcy);
    return hipGetLastError();
}

template <typename T>
hipError_t Stages<T>::halo_xy(const Geo& g, T* a, int mode, int full, hipStream_t s) {
    // y walls first, then x walls (mpdataoperators.inc:487-511)
    const int ext = full ? 2 * g.ih : 0;
    if (!g.py) {
        dim3 grid((g.np + ext + 63) / 64, (g.lp + ext + 3) / 4, 2);
        k_halo_wall<T><<<grid, wall_block(), 0, s>>>(g, a, 1, mode, full);
    }
    if (!g.px) {
        dim3 grid((g.mp + ext + 63) / 64, (g.lp + ext + 3) / 4, 2);
        k_halo_wall<T><<<grid, wall_block(), 0, s>>>(g, a, 0, mode, full);
    }
    return hipGetLastError();
}

template <typename T>
hipError_t Stages<T>::face_bc(const Geo& g, T* a1, T* a2, T* a3, hipStream_t s) {
    if ((a1 && !g.px) || (a2 && !g.py)) {
        int a = (g.mp > g.np ? g.mp : g.np) + 2 * g.ih;
        dim3 grid((a + 63) / 64, (g.lp + 2 * g.ih + 3) / 4, 4);
        k_face_bc_xy<T><<<grid, wall_block(), 0, s>>>(g, a1, a2);
    }
    if (a3 && !g.pz) {
        dim3 grid((g.np + 63) / 64, (g.mp + 3) / 4, 1);
        k_face_bc_z<T><<<grid, wall_block(), 0, s>>>(g, a3);
    }
    return hipGetLastError();
}

template <typename T>
hipError_t Stages<T>::upwind(const Geo& g, const T* u1, const T* u2, const T* u3, const T* x,
                             T* xant, const T* rhr, const T* h, hipStream_t s) {
    k_upwind<T><<<cell_grid(g), cell_block(), 0, s>>>(g, u1, u2, u3, x, xant, rhr, h);
    return hipGetLastError();
}

template <typename T>
hipError_t Stages<T>::pseudo_vel(const Geo& g, int variant, const T* u1, const T* u2,
                                 const T* u3, const T* xa, const T* h, T* v1, T* v2, T* v3,
                                 T* f1ns, T* f2ns, T* f3ns, hipStream_t s) {
    if (variant == 1)
        k_pseudo_vel<T, 1><<<cell_grid(g), cell_block(), 0, s>>>(g, u1, u2, u3, xa, h, v1, v2, v3,
                                                                 f1ns, f2ns, f3ns);
    else
        k_pseudo_vel<T, 2><<<cell_grid(g), cell_block(), 0, s>>>(g, u1, u2, u3, xa, h, v1, v2, v3,
                                                                 f1ns, f2ns, f3ns);
    return hipGetLastError();
}

template <typename T>
hipError_t Stages<T>::limiter(const Geo& g, const T* x0, const T* xa, const T* h, const T* w1,
                              const T* w2, const T* w3, T* cp, T* cn, hipStream_t s) {
    k_limiter<T><<<cell_grid(g), cell_block(), 0, s>>>(g, x0, xa, h, w1, w2, w3, cp, cn);
    return hipGetLastError();
}

template <typename T>
hipError_t Stages<T>::gauge_update(const Geo& g, int nonos, const T* v1, const T* v2,
                                   const T* v3, const T* cp, const T* cn, const T* psi,
                                   const T* rhr, const T* h, T* out, T* ul1, T* ul2, T* ul3,
                                   hipStream_t s) {
    if (nonos)
        k_gauge_update<T, 1><<<cell_grid(g), cell_block(), 0, s>>>(g, v1, v2, v3, cp, cn, psi, rhr,
                                                                   h, out, ul1, ul2, ul3);
    else
        k_gauge_update<T, 0><<<cell_grid(g), cell_block(), 0, s>>>(g, v1, v2, v3, cp, cn, psi, rhr,
                                                                   h, out, ul1, ul2, ul3);
    return hipGetLastError();
}

template <typename T>
hipError_t Stages<T>::std_flux(const Geo& g, int nonos, const T* v1, const T* v2, const T* v3,
                               const T* cp, const T* cn, const T* psi, const T* rhr, T* f1, T* f2,
                               T* f3, T* ul1, T* ul2, T* ul3, hipStream_t s) {
    if (nonos)
        k_std_flux<T, 1><<<cell_grid(g), cell_block(), 0, s>>>(g, v1, v2, v3, cp, cn, psi, rhr, f1,
                                                               f2, f3, ul1, ul2, ul3);
    else
        k_std_flux<T, 0><<<cell_grid(g), cell_block(), 0, s>>>(g, v1, v2, v3, cp, cn, psi, rhr, f1,
                                                               f2, f3, ul1, ul2, ul3);
    return hipGetLastError();
}

template <typename T>
hipError_t Stages<T>::std_update(const Geo& g, const T* f1, const T* f2, const T* f3,
                                 const T* psi, const T* rhr, const T* h, T* out, hipStream_t s) {
    k_std_update<T><<<cell_grid(g), cell_block(), 0, s>>>(g, f1, f2, f3, psi, rhr, h, out);
    return hipGetLastError();
}

template <typename T>
hipError_t Stages<T>::copy_interior(const Geo& g, const T* src, T* dst, hipStream_t s) {
    k_copy_interior<T><<<cell_grid(g), cell_block(), 0, s>>>(g, src, dst);
    return hipGetLastError();
}

template <typename T>
hipError_t Stages<T>::velprd_extrap(const Geo& g, T* ox, T* oy, T* oz, const T* h, T gc1, T gc2,
                                    T gc3, T dt_ratio, hipStream_t s) {
    k_velprd_extrap<T><<<cell_grid(g), cell_block(), 0, s>>>(g, ox, oy, oz, h, gc1, gc2, gc3,
                                                             dt_ratio);
    return hipGetLastError();
}

template <typename T>
hipError_t Stages<T>::a_to_c(const Geo& g, T* ox, T* oy, T* oz, T* u1, T* u2, T* u3,
                             hipStream_t s) {
    if (!g.px) {
        dim3 grid((g.mp + 63) / 64, (g.lp + 3) / 4, 2);
        k_extrap_halo<T><<<grid, wall_block(), 0, s>>>(g, ox, 0);
    }
    if (!g.py) {
        dim3 grid((g.np + 63) / 64, (g.lp + 3) / 4, 2);
        k_extrap_halo<T><<<grid, wall_block(), 0, s>>>(g, oy, 1);
    }
    if (!g.pz) {
        dim3 grid((g.np + 63) / 64, (g.mp + 3) / 4, 2);
        k_extrap_halo<T><<<grid, wall_block(), 0, s>>>(g, oz, 2);
    }
    k_a_to_c<T><<<cell_grid(g, 1), cell_block(), 0, s>>>(g, ox, oy, oz, u1, u2, u3);
    return hipGetLastError();
}

template <typename T>
hipError_t Stages<T>::predict_ac(const Geo& g, const T* ox, const T* oy, const T* oz,
                                 const T* oxp, const T* oyp, const T* ozp, const T* h, T gc1,
                                 T gc2, T gc3, T dt_ratio, T* u1, T* u2, T* u3, hipStream_t s) {
    k_predict_ac<T><<<cell_grid(g, 1), cell_block(), 0, s>>>(g, ox, oy, oz, oxp, oyp, ozp, h, gc1,
                                                             gc2, gc3, dt_ratio, u1, u2, u3);
    return hipGetLastError();
}

template <typename T>
hipError_t Stages<T>::velprd_rotate(const Geo& g, T* u, T* v, T* w, T* ox, T* oy, T* oz,
                                    const T* h, T r, hipStream_t s) {
    k_velprd_rotate<T><<<cell_grid(g), cell_block(), 0, s>>>(g, u, v, w, ox, oy, oz, h, r);
    return hipGetLastError();
}

template <typename T>
long long Stages<T>::reduce_blocks(const Geo& g) {
    dim3 gr = cell_grid(g);
    return (long long)gr.x * gr.y * gr.z;
}

template <typename T>
hipError_t Stages<T>::courants(const Geo& g, const T* ox, const T* oy, const T* oz, const T* h,
                               int level, T gc1, T gc2, T gc3, T* part, hipStream_t s) {
    k_courants<T><<<cell_grid(g), cell_block(), 0, s>>>(g, ox, oy, oz, h, level, gc1, gc2, gc3,
                                                        part);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_max7<T><<<1, 256, 0, s>>>(part, reduce_blocks(g), part + 7 * reduce_blocks(g));
    return hipGetLastError();
}

template <typename T>
hipError_t Stages<T>::norm_sums(const Geo& g, const T* x, const T* x0, double* part,
                                hipStream_t s) {
    k_norm_sums<T><<<cell_grid(g), cell_block(), 0, s>>>(g, x, x0, part);
    return hipGetLastError();
}

template struct Stages<double>;
template struct Stages<float>;

// The stages of one step on (value, tangent) pairs: the tangent-linear step.
using D64 = Dual<double>;
template hipError_t Stages<D64>::putbcinxh(const Geo&, D64*, const D64*, const D64*, hipStream_t);
template hipError_t Stages<D64>::halo_xy(const Geo&, D64*, int, int, hipStream_t);
template hipError_t Stages<D64>::face_bc(const Geo&, D64*, D64*, D64*, hipStream_t);
template hipError_t Stages<D64>::upwind(const Geo&, const D64*, const D64*, const D64*,
                                        const D64*, D64*, const D64*, const D64*, hipStream_t);
template hipError_t Stages<D64>::pseudo_vel(const Geo&, int, const D64*, const D64*, const D64*,
                                            const D64*, const D64*, D64*, D64*, D64*, D64*, D64*,
                                            D64*, hipStream_t);
template hipError_t Stages<D64>::limiter(const Geo&, const D64*, const D64*, const D64*,
                                         const D64*, const D64*, const D64*, D64*, D64*,
                                         hipStream_t);
template hipError_t Stages<D64>::gauge_update(const Geo&, int, const D64*, const D64*, const D64*,
                                              const D64*, const D64*, const D64*, const D64*,
                                              const D64*, D64*, D64*, D64*, D64*, hipStream_t);
template hipError_t Stages<D64>::std_flux(const Geo&, int, const D64*, const D64*, const D64*,
                                          const D64*, const D64*, const D64*, const D64*, D64*,
                                          D64*, D64*, D64*, D64*, D64*, hipStream_t);
template hipError_t Stages<D64>::std_update(const Geo&, const D64*, const D64*, const D64*,
                                            const D64*, const D64*, const D64*, D64*, hipStream_t);
template hipError_t Stages<D64>::copy_interior(const Geo&, const D64*, D64*, hipStream_t);

}  // namespace mpdata
