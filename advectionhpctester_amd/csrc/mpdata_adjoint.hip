// Adjoint of the standard non-oscillatory MPDATA step (2 iterations, open
// walls, FP64): a recorded forward step keeps its intermediates in a tape
// slot (mpdata_adj_record_f64), and mpdata_adj_step_f64 runs the reverse of
// each stage of antidiff3d-standard-gpubc.f90 and upwind3d-gpubc.f90 in
// reverse order:
//   update (:431-459) -> limited donor fluxes (:372-407) -> limiter
//   (:260-347) -> pseudo-velocities and first-order fluxes (:72-202) ->
//   halo fills (mirror of xant, edge copy of x) -> upwind (:101-198).
// Each reverse kernel transposes the linearisation the tangent-linear step
// (the stage kernels on dual numbers) computes: every branch is taken as the
// forward value takes it.  Contributions from one face or cell to several
// neighbours are scattered with atomic adds into the adjoint arrays, so the
// adjoint is deterministic up to the order of those additions.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/mpdata_hip.h"
#include "mpdata_ctx.hpp"
#include "mpdata_stages.hpp"
#include "mpdata_status.hpp"

using namespace mpdata;

namespace {

#define CELL_HEAD                                                  \
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;       \
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;       \
    const int k = 1 + blockIdx.z;                                  \
    if (i > g.np || j > g.mp || k > g.lp) return;                  \
    const bool gnd = k == 1, sky = k == g.lp;

#define S_(a, ii, jj, kk) a[off_raw<SCAL>(g, ii, jj, kk)]
#define X_(a, ii, jj, kk) a[off_raw<FACEX>(g, ii, jj, kk)]
#define Y_(a, ii, jj, kk) a[off_raw<FACEY>(g, ii, jj, kk)]
#define Z_(a, ii, jj, kk) a[off_raw<FACEZ>(g, ii, jj, kk)]
#define SP(a, ii, jj, kk) (&a[off_raw<SCAL>(g, ii, jj, kk)])
#define XP(a, ii, jj, kk) (&a[off_raw<FACEX>(g, ii, jj, kk)])
#define YP(a, ii, jj, kk) (&a[off_raw<FACEY>(g, ii, jj, kk)])
#define ZP(a, ii, jj, kk) (&a[off_raw<FACEZ>(g, ii, jj, kk)])

using R = double;

__device__ __forceinline__ R ind_pos(R y) { return y > R(0) ? R(1) : R(0); }  // d pp / dy
__device__ __forceinline__ R ind_neg(R y) { return y < R(0) ? R(1) : R(0); }  // d mn0 / dy
__device__ __forceinline__ R sgn(R y) { return y > R(0) ? R(1) : (y < R(0) ? R(-1) : R(0)); }
__device__ __forceinline__ void acc(R* p, R v) { atomicAdd(p, v); }

// adjoint work arrays held in the context (index into mpdata_ctx::adj)
enum { A_PSI = 0, A_X0, A_CP, A_CN, A_SUM, A_V1, A_V2, A_V3, A_F1, A_F2, A_F3, A_N };
const int A_KIND[A_N] = {SCAL, SCAL, SCAL, SCAL, SCAL, FACEX, FACEY, FACEZ, FACEX, FACEY, FACEZ};

struct Adj {
    R *psi, *x0, *cp, *cn, *sum, *v1, *v2, *v3, *f1, *f2, *f3;
};

__global__ void k_zero(R* __restrict__ a, long long n) {
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
         q += (long long)gridDim.x * blockDim.x)
        a[q] = R(0);
}

// ---- update: out = rhr*(psi/rhr - sum*hinv) (antidiff3d-standard-gpubc.f90:431-459)
__global__ void __launch_bounds__(256) k_adj_update(Geo g, const R* __restrict__ xbar,
                                                    const R* __restrict__ rhr,
                                                    const R* __restrict__ h, Adj a) {
    CELL_HEAD
    (void)gnd; (void)sky;
    const R w = S_(rhr, i, j, k) * S_(xbar, i, j, k);
    const R hinv = R(1) / S_(h, i, j, k);
    S_(a.psi, i, j, k) = w / S_(rhr, i, j, k);
    S_(a.sum, i, j, k) = -(w * hinv);
}

// ---- limited donor fluxes (:372-407), one face ----------------------------
// fb = adjoint of the flux; m/c = upstream/downstream cells
__device__ __forceinline__ void adj_face_flux(R fb, R v, R vl, R pm, R pc, R rm, R rc,
                                              R cp_c, R cn_m, R cp_m, R cn_c, R* psib_m,
                                              R* psib_c, R* cpb_c, R* cnb_m, R* cpb_m,
                                              R* cnb_c, R* vbar) {
    const R y1 = pm / rm, y2 = pc / rc;
    // f = donor(y1, y2, vl)
    const R y1b = fb * (vl > R(0) ? vl : R(0));
    const R y2b = fb * (vl < R(0) ? vl : R(0));
    const R vlb = fb * (ind_pos(vl) * y1 + ind_neg(vl) * y2);
    acc(psib_m, y1b / rm);
    acc(psib_c, y2b / rc);
    // vl = slimv(v, a1, a2, pm, pc)
    const R Pm = copysign(R(1), pm) > R(0) ? R(1) : R(0);
    const R Nm = copysign(R(1), -pm) > R(0) ? R(1) : R(0);
    const R Pc = copysign(R(1), pc) > R(0) ? R(1) : R(0);
    const R Nc = copysign(R(1), -pc) > R(0) ? R(1) : R(0);
    const R m1 = R(1) < cp_c ? R(1) : cp_c;
    const R a1 = m1 < cn_m ? m1 : cn_m;
    const R m2 = R(1) < cp_m ? R(1) : cp_m;
    const R a2 = m2 < cn_c ? m2 : cn_c;
    const R ppv = v > R(0) ? v : R(0), pnv = v < R(0) ? -v : R(0);
    const R K1 = a1 * Pm + a2 * Nm, K2 = a2 * Pc + a1 * Nc;
    *vbar = vlb * (ind_pos(v) * K1 + ind_neg(v) * K2);
    const R a1b = vlb * (ppv * Pm - pnv * Nc);
    const R a2b = vlb * (ppv * Nm - pnv * Pc);
    // a1 = min3(1, cp_c, cn_m), a2 = min3(1, cp_m, cn_c)
    if (m1 < cn_m) {
        if (!(R(1) < cp_c)) acc(cpb_c, a1b);
    } else {
        acc(cnb_m, a1b);
    }
    if (m2 < cn_c) {
        if (!(R(1) < cp_m)) acc(cpb_m, a2b);
    } else {
        acc(cnb_c, a2b);
    }
}

__global__ void __launch_bounds__(256) k_adj_flux(Geo g, const R* __restrict__ v1,
                                                  const R* __restrict__ v2,
                                                  const R* __restrict__ v3,
                                                  const R* __restrict__ vl1,
                                                  const R* __restrict__ vl2,
                                                  const R* __restrict__ vl3,
                                                  const R* __restrict__ cp,
                                                  const R* __restrict__ cn,
                                                  const R* __restrict__ psi,
                                                  const R* __restrict__ rhr, Adj a) {
    CELL_HEAD
    const R pc = S_(psi, i, j, k), rc = S_(rhr, i, j, k);
    const R sc = S_(a.sum, i, j, k);
    if (i > 1) {
        const R fb = S_(a.sum, i - 1, j, k) - sc;
        adj_face_flux(fb, X_(v1, i, j, k), X_(vl1, i, j, k), S_(psi, i - 1, j, k), pc,
                      S_(rhr, i - 1, j, k), rc, S_(cp, i, j, k), S_(cn, i - 1, j, k),
                      S_(cp, i - 1, j, k), S_(cn, i, j, k), SP(a.psi, i - 1, j, k),
                      SP(a.psi, i, j, k), SP(a.cp, i, j, k), SP(a.cn, i - 1, j, k),
                      SP(a.cp, i - 1, j, k), SP(a.cn, i, j, k), XP(a.v1, i, j, k));
    }
    if (j > 1) {
        const R fb = S_(a.sum, i, j - 1, k) - sc;
        adj_face_flux(fb, Y_(v2, i, j, k), Y_(vl2, i, j, k), S_(psi, i, j - 1, k), pc,
                      S_(rhr, i, j - 1, k), rc, S_(cp, i, j, k), S_(cn, i, j - 1, k),
                      S_(cp, i, j - 1, k), S_(cn, i, j, k), SP(a.psi, i, j - 1, k),
                      SP(a.psi, i, j, k), SP(a.cp, i, j, k), SP(a.cn, i, j - 1, k),
                      SP(a.cp, i, j - 1, k), SP(a.cn, i, j, k), YP(a.v2, i, j, k));
    }
    if (k > 1) {
        // the ground cell counts its upper face twice, the sky cell its lower
        const R c1 = k - 1 == 1 ? R(2) : R(1), c2 = sky ? R(2) : R(1);
        const R fb = c1 * S_(a.sum, i, j, k - 1) - c2 * sc;
        adj_face_flux(fb, Z_(v3, i, j, k), Z_(vl3, i, j, k), S_(psi, i, j, k - 1), pc,
                      S_(rhr, i, j, k - 1), rc, S_(cp, i, j, k), S_(cn, i, j, k - 1),
                      S_(cp, i, j, k - 1), S_(cn, i, j, k), SP(a.psi, i, j, k - 1),
                      SP(a.psi, i, j, k), SP(a.cp, i, j, k), SP(a.cn, i, j, k - 1),
                      SP(a.cp, i, j, k - 1), SP(a.cn, i, j, k), ZP(a.v3, i, j, k));
    }
    (void)gnd;
}

// ---- limiter (:260-347) ----------------------------------------------------
// max/min over x0 and psi with the forward kernel's order of comparisons;
// `which` codes the selected operand: 0 = x0, 1 = psi, with its cell.
struct Pick {
    R val;
    int src, di, dj, dk;
};
__device__ __forceinline__ void pick_max(Pick& p, R v, int src, int di, int dj, int dk) {
    if (!(p.val > v)) p = Pick{v, src, di, dj, dk};   // max2(hi, t) = hi > t ? hi : t
}
__device__ __forceinline__ void pick_min(Pick& p, R v, int src, int di, int dj, int dk) {
    if (!(p.val < v)) p = Pick{v, src, di, dj, dk};   // min2(lo, t) = lo < t ? lo : t
}

__global__ void __launch_bounds__(256) k_adj_limiter(Geo g, const R* __restrict__ x0,
                                                     const R* __restrict__ psi,
                                                     const R* __restrict__ h,
                                                     const R* __restrict__ w1,
                                                     const R* __restrict__ w2,
                                                     const R* __restrict__ w3,
                                                     const R* __restrict__ cp,
                                                     const R* __restrict__ cn, Adj a) {
    CELL_HEAD
    const R ep = Eps<R>::v;
    Pick mxo{S_(x0, i - 1, j, k), 0, -1, 0, 0}, mno = mxo;
    Pick mx{S_(psi, i - 1, j, k), 1, -1, 0, 0}, mn = mx;
#define ACC(arr, src, lo, hi, di, dj, dk)                      \
    {                                                          \
        const R t_ = S_(arr, i + di, j + dj, k + dk);          \
        pick_min(lo, t_, src, di, dj, dk);                     \
        pick_max(hi, t_, src, di, dj, dk);                     \
    }
    ACC(x0, 0, mno, mxo, 0, 0, 0) ACC(x0, 0, mno, mxo, 1, 0, 0)
    ACC(x0, 0, mno, mxo, 0, -1, 0) ACC(x0, 0, mno, mxo, 0, 1, 0)
    ACC(psi, 1, mn, mx, 0, 0, 0) ACC(psi, 1, mn, mx, 1, 0, 0)
    ACC(psi, 1, mn, mx, 0, -1, 0) ACC(psi, 1, mn, mx, 0, 1, 0)
    if (!gnd) { ACC(x0, 0, mno, mxo, 0, 0, -1) ACC(psi, 1, mn, mx, 0, 0, -1) }
    if (!sky) { ACC(x0, 0, mno, mxo, 0, 0, 1) ACC(psi, 1, mn, mx, 0, 0, 1) }
#undef ACC
    if (!(mx.val > mxo.val)) mx = mxo;   // mx = max2(mx, mxo)
    if (!(mn.val < mno.val)) mn = mno;   // mn = min2(mn, mno)
    const R a1p = X_(w1, i + 1, j, k), a1 = X_(w1, i, j, k);
    const R a2p = Y_(w2, i, j + 1, k), a2 = Y_(w2, i, j, k);
    R a3p, a3;
    if (gnd) { a3p = Z_(w3, i, j, k + 1); a3 = -a3p; }
    else if (sky) { a3 = Z_(w3, i, j, k); a3p = -a3; }
    else { a3p = Z_(w3, i, j, k + 1); a3 = Z_(w3, i, j, k); }
    const R hc = S_(h, i, j, k);
    const R qp = (a1p < R(0) ? -a1p : R(0)) + (a1 > R(0) ? a1 : R(0)) + (a2p < R(0) ? -a2p : R(0)) +
                 (a2 > R(0) ? a2 : R(0)) + (a3p < R(0) ? -a3p : R(0)) + (a3 > R(0) ? a3 : R(0)) + ep;
    const R qn = (a1p > R(0) ? a1p : R(0)) + (a1 < R(0) ? -a1 : R(0)) + (a2p > R(0) ? a2p : R(0)) +
                 (a2 < R(0) ? -a2 : R(0)) + (a3p > R(0) ? a3p : R(0)) + (a3 < R(0) ? -a3 : R(0)) + ep;
    const R cpb = S_(a.cp, i, j, k), cnb = S_(a.cn, i, j, k);
    const R tpb = cpb / qp, tnb = cnb / qn;
    const R qpb = -(cpb * S_(cp, i, j, k) / qp);
    const R qnb = -(cnb * S_(cn, i, j, k) / qn);
    // cp = (mx - xc)*hc/qp, cn = (xc - mn)*hc/qn
    const R mxb = tpb * hc, mnb = -(tnb * hc);
    acc(SP(a.psi, i, j, k), -(tpb * hc) + tnb * hc);
    acc(mx.src ? SP(a.psi, i + mx.di, j + mx.dj, k + mx.dk) : SP(a.x0, i + mx.di, j + mx.dj, k + mx.dk), mxb);
    acc(mn.src ? SP(a.psi, i + mn.di, j + mn.dj, k + mn.dk) : SP(a.x0, i + mn.di, j + mn.dj, k + mn.dk), mnb);
    // denominators: d pn(y) = -[y<0] dy, d pp(y) = [y>0] dy
    const R b1p = -ind_neg(a1p) * qpb + ind_pos(a1p) * qnb;  // a1p enters qp as pn, qn as pp
    const R b1 = ind_pos(a1) * qpb - ind_neg(a1) * qnb;
    const R b2p = -ind_neg(a2p) * qpb + ind_pos(a2p) * qnb;
    const R b2 = ind_pos(a2) * qpb - ind_neg(a2) * qnb;
    const R b3p = -ind_neg(a3p) * qpb + ind_pos(a3p) * qnb;
    const R b3 = ind_pos(a3) * qpb - ind_neg(a3) * qnb;
    // wall faces of x and y are constants after the boundary condition
    if (i < g.np) acc(XP(a.f1, i + 1, j, k), b1p);
    if (i > 1) acc(XP(a.f1, i, j, k), b1);
    if (j < g.mp) acc(YP(a.f2, i, j + 1, k), b2p);
    if (j > 1) acc(YP(a.f2, i, j, k), b2);
    if (gnd) acc(ZP(a.f3, i, j, k + 1), b3p - b3);
    else if (sky) acc(ZP(a.f3, i, j, k), b3 - b3p);
    else {
        acc(ZP(a.f3, i, j, k + 1), b3p);
        acc(ZP(a.f3, i, j, k), b3);
    }
}

// ---- pseudo-velocities and first-order fluxes (:72-202) --------------------
// V = vdyfa(x1, x2, a, r) + sum of vcorra(a, b, y1, y2, r) terms; each vcorra
// term carries its four b faces and the four cells of its y1/y2 sums.
struct Corr {
    R b, y1, y2;
    const R* bf[4];   // b = sum of these faces
    R* bfb[4];        // their adjoints
    R* cb[4];         // adjoints of the cells (plus pair: cells 0,1; minus pair: 2,3)
    R cv[4];          // cell values
};

__device__ __forceinline__ void adj_vcorra(R vb, R a, const Corr& c, R r, R* ab) {
    // val = -((((0.125*a)*b)*y1)*r / y2)
    const R t1 = R(0.125) * a, t2 = t1 * c.b, t3 = t2 * c.y1, t4 = t3 * r, q = t4 / c.y2;
    const R qb = -vb;
    const R t4b = qb / c.y2;
    const R y2b = -(qb * q / c.y2);
    const R t3b = t4b * r;
    const R y1b = t3b * t2;
    const R t2b = t3b * c.y1;
    const R bb = t2b * t1;
    const R t1b = t2b * c.b;
    *ab += R(0.125) * t1b;
    for (int q2 = 0; q2 < 4; q2++) acc(c.bfb[q2], bb);
    // y1 = (|c0| + |c1|) - (|c2| + |c3|), y2 = |c0| + |c1| + |c2| + |c3| + ep
    acc(c.cb[0], (y1b + y2b) * sgn(c.cv[0]));
    acc(c.cb[1], (y1b + y2b) * sgn(c.cv[1]));
    acc(c.cb[2], (-y1b + y2b) * sgn(c.cv[2]));
    acc(c.cb[3], (-y1b + y2b) * sgn(c.cv[3]));
}

// vb: adjoint of V at this face; fb: adjoint of fns = donor(x1, x2, V)
__device__ __forceinline__ void adj_face_pseudo(R vb, R fb, R V, R x1, R x2, R a, R r, R* x1b,
                                                R* x2b, R* ab_out, const Corr* cs, int ncorr) {
    const R ep = Eps<R>::v;
    // fns = donor(x1, x2, V)
    vb += fb * (ind_pos(V) * x1 + ind_neg(V) * x2);
    R x1bv = fb * (V > R(0) ? V : R(0));
    R x2bv = fb * (V < R(0) ? V : R(0));
    // vdyfa: (|a| - a*a*r) * (|x2| - |x1|) / (|x2| + |x1| + ep)
    const R P = fabs(a) - a * a * r, N = fabs(x2) - fabs(x1), D = fabs(x2) + fabs(x1) + ep;
    const R val = P * N / D;
    const R tb = vb / D, Db = -(vb * val / D);
    const R Pb = tb * N, Nb = tb * P;
    x2bv += (Nb + Db) * sgn(x2);
    x1bv += (-Nb + Db) * sgn(x1);
    R ab = Pb * sgn(a) - Pb * r * (a + a);
    for (int q = 0; q < ncorr; q++) adj_vcorra(vb, a, cs[q], r, &ab);
    acc(x1b, x1bv);
    acc(x2b, x2bv);
    acc(ab_out, ab);
}

__global__ void __launch_bounds__(256) k_adj_pseudo(Geo g, const R* __restrict__ u1,
                                                    const R* __restrict__ u2,
                                                    const R* __restrict__ u3,
                                                    const R* __restrict__ psi,
                                                    const R* __restrict__ h,
                                                    const R* __restrict__ v1,
                                                    const R* __restrict__ v2,
                                                    const R* __restrict__ v3, R* __restrict__ ub1,
                                                    R* __restrict__ ub2, R* __restrict__ ub3, Adj a) {
    CELL_HEAD
    const bool zterms = !(gnd || sky);
    const R ep = Eps<R>::v;
#define AXV(ii, jj, kk) S_(psi, ii, jj, kk)
#define CORR(c, F1, F2, F3, F4, BF1, BF2, BF3, BF4, I0, J0, K0, I1, J1, K1, I2, J2, K2, I3, J3, K3) \
    {                                                                                          \
        c.bf[0] = &F1; c.bf[1] = &F2; c.bf[2] = &F3; c.bf[3] = &F4;                            \
        c.bfb[0] = BF1; c.bfb[1] = BF2; c.bfb[2] = BF3; c.bfb[3] = BF4;                        \
        c.b = *c.bf[0] + *c.bf[1] + *c.bf[2] + *c.bf[3];                                       \
        c.cv[0] = AXV(I0, J0, K0); c.cv[1] = AXV(I1, J1, K1);                                  \
        c.cv[2] = AXV(I2, J2, K2); c.cv[3] = AXV(I3, J3, K3);                                  \
        c.cb[0] = SP(a.psi, I0, J0, K0); c.cb[1] = SP(a.psi, I1, J1, K1);                      \
        c.cb[2] = SP(a.psi, I2, J2, K2); c.cb[3] = SP(a.psi, I3, J3, K3);                      \
        c.y1 = (fabs(c.cv[0]) + fabs(c.cv[1])) - (fabs(c.cv[2]) + fabs(c.cv[3]));              \
        c.y2 = fabs(c.cv[0]) + fabs(c.cv[1]) + fabs(c.cv[2]) + fabs(c.cv[3]) + ep;             \
    }
    if (i > 1) {
        const R hm = R(1) / (R(0.5) * (S_(h, i - 1, j, k) + S_(h, i, j, k)));
        Corr cs[2];
        CORR(cs[0], Y_(u2, i - 1, j, k), Y_(u2, i - 1, j + 1, k), Y_(u2, i, j + 1, k), Y_(u2, i, j, k),
             YP(ub2, i - 1, j, k), YP(ub2, i - 1, j + 1, k), YP(ub2, i, j + 1, k), YP(ub2, i, j, k),
             i - 1, j + 1, k, i, j + 1, k, i - 1, j - 1, k, i, j - 1, k)
        if (zterms)
            CORR(cs[1], Z_(u3, i - 1, j, k), Z_(u3, i - 1, j, k + 1), Z_(u3, i, j, k + 1), Z_(u3, i, j, k),
                 ZP(ub3, i - 1, j, k), ZP(ub3, i - 1, j, k + 1), ZP(ub3, i, j, k + 1), ZP(ub3, i, j, k),
                 i - 1, j, k + 1, i, j, k + 1, i - 1, j, k - 1, i, j, k - 1)
        adj_face_pseudo(X_(a.v1, i, j, k), X_(a.f1, i, j, k), X_(v1, i, j, k), AXV(i - 1, j, k),
                        AXV(i, j, k), X_(u1, i, j, k), hm, SP(a.psi, i - 1, j, k),
                        SP(a.psi, i, j, k), XP(ub1, i, j, k), cs, zterms ? 2 : 1);
    }
    if (j > 1) {
        const R hm = R(1) / (R(0.5) * (S_(h, i, j - 1, k) + S_(h, i, j, k)));
        Corr cs[2];
        CORR(cs[0], X_(u1, i, j - 1, k), X_(u1, i, j, k), X_(u1, i + 1, j, k), X_(u1, i + 1, j - 1, k),
             XP(ub1, i, j - 1, k), XP(ub1, i, j, k), XP(ub1, i + 1, j, k), XP(ub1, i + 1, j - 1, k),
             i + 1, j - 1, k, i + 1, j, k, i - 1, j - 1, k, i - 1, j, k)
        if (zterms)
            CORR(cs[1], Z_(u3, i, j - 1, k), Z_(u3, i, j, k), Z_(u3, i, j, k + 1), Z_(u3, i, j - 1, k + 1),
                 ZP(ub3, i, j - 1, k), ZP(ub3, i, j, k), ZP(ub3, i, j, k + 1), ZP(ub3, i, j - 1, k + 1),
                 i, j - 1, k + 1, i, j, k + 1, i, j - 1, k - 1, i, j, k - 1)
        adj_face_pseudo(Y_(a.v2, i, j, k), Y_(a.f2, i, j, k), Y_(v2, i, j, k), AXV(i, j - 1, k),
                        AXV(i, j, k), Y_(u2, i, j, k), hm, SP(a.psi, i, j - 1, k),
                        SP(a.psi, i, j, k), YP(ub2, i, j, k), cs, zterms ? 2 : 1);
    }
    if (!gnd) {
        const R hm = R(1) / (R(0.5) * (S_(h, i, j, k - 1) + S_(h, i, j, k)));
        Corr cs[2];
        CORR(cs[0], X_(u1, i, j, k - 1), X_(u1, i, j, k), X_(u1, i + 1, j, k), X_(u1, i + 1, j, k - 1),
             XP(ub1, i, j, k - 1), XP(ub1, i, j, k), XP(ub1, i + 1, j, k), XP(ub1, i + 1, j, k - 1),
             i + 1, j, k - 1, i + 1, j, k, i - 1, j, k - 1, i - 1, j, k)
        CORR(cs[1], Y_(u2, i, j, k - 1), Y_(u2, i, j + 1, k - 1), Y_(u2, i, j + 1, k), Y_(u2, i, j, k),
             YP(ub2, i, j, k - 1), YP(ub2, i, j + 1, k - 1), YP(ub2, i, j + 1, k), YP(ub2, i, j, k),
             i, j + 1, k - 1, i, j + 1, k, i, j - 1, k - 1, i, j - 1, k)
        adj_face_pseudo(Z_(a.v3, i, j, k), Z_(a.f3, i, j, k), Z_(v3, i, j, k), AXV(i, j, k - 1),
                        AXV(i, j, k), Z_(u3, i, j, k), hm, SP(a.psi, i, j, k - 1),
                        SP(a.psi, i, j, k), ZP(ub3, i, j, k), cs, 2);
    }
#undef CORR
#undef AXV
}

// ---- halo fills in reverse -------------------------------------------------
// x direction (dir 0): the mirror of xant over the full (j,k) range and the
// edge copy of x over the interior range; y direction (dir 1) likewise.  The
// forward fills y first, then x, so the reverse runs dir 0, then dir 1.
__global__ void k_adj_halo(Geo g, int dir, Adj a) {
    const int ih = g.ih;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;   // along the wall
    const int k = blockIdx.y + 1 - ih;                     // 1-ih .. lp+ih
    if (dir == 0) {
        const int j = p + 1 - ih;
        if (j > g.mp + ih) return;
        R* lo = SP(a.psi, 0, j, k);
        R* hi = SP(a.psi, g.np + 1, j, k);
        *SP(a.psi, 2, j, k) += *lo;
        *lo = R(0);
        *SP(a.psi, g.np - 1, j, k) += *hi;
        *hi = R(0);
        if (j >= 1 && j <= g.mp && k >= 1 && k <= g.lp) {
            *SP(a.x0, 1, j, k) += *SP(a.x0, 0, j, k);
            *SP(a.x0, g.np, j, k) += *SP(a.x0, g.np + 1, j, k);
        }
    } else {
        const int i = p + 1 - ih;
        if (i > g.np + ih) return;
        R* lo = SP(a.psi, i, 0, k);
        R* hi = SP(a.psi, i, g.mp + 1, k);
        *SP(a.psi, i, 2, k) += *lo;
        *lo = R(0);
        *SP(a.psi, i, g.mp - 1, k) += *hi;
        *hi = R(0);
        if (i >= 1 && i <= g.np && k >= 1 && k <= g.lp) {
            *SP(a.x0, i, 1, k) += *SP(a.x0, i, 0, k);
            *SP(a.x0, i, g.mp, k) += *SP(a.x0, i, g.mp + 1, k);
        }
    }
}

// ---- upwind (upwind3d-gpubc.f90:101-198) and the sum of both uses of x -----
// xant = rhr*(x - sum*hinv), fluxes donor(x(m), x(c), u); the x halo holds the
// inflow values (no adjoint).  Each cell takes its low faces (and the last
// cell of a direction its high wall face) for the velocity adjoint.
__device__ __forceinline__ R sbar_of(Geo g, const Adj& a, const R* rhr, const R* h, int i, int j,
                                     int k) {
    return -(S_(rhr, i, j, k) * S_(a.psi, i, j, k) * (R(1) / S_(h, i, j, k)));
}

__global__ void __launch_bounds__(256) k_adj_upwind(Geo g, const R* __restrict__ u1,
                                                    const R* __restrict__ u2,
                                                    const R* __restrict__ u3,
                                                    const R* __restrict__ xb,
                                                    const R* __restrict__ rhr,
                                                    const R* __restrict__ h, R* __restrict__ xbar,
                                                    R* __restrict__ ub1, R* __restrict__ ub2,
                                                    R* __restrict__ ub3, Adj a) {
    CELL_HEAD
    const int np = g.np, mp = g.mp, lp = g.lp;
    const R w = S_(rhr, i, j, k) * S_(a.psi, i, j, k);
    const R sc = -(w * (R(1) / S_(h, i, j, k)));
    R xb_c = w + S_(a.x0, i, j, k);
    const R xc = S_(xb, i, j, k);
    // x faces i (c downstream) and i+1 (c upstream)
    {
        const R fl = (i > 1 ? sbar_of(g, a, rhr, h, i - 1, j, k) : R(0)) - sc;
        const R fh = sc - (i < np ? sbar_of(g, a, rhr, h, i + 1, j, k) : R(0));
        const R ul = X_(u1, i, j, k), uh = X_(u1, i + 1, j, k);
        xb_c += fl * (ul < R(0) ? ul : R(0)) + fh * (uh > R(0) ? uh : R(0));
        X_(ub1, i, j, k) += fl * (ind_pos(ul) * S_(xb, i - 1, j, k) + ind_neg(ul) * xc);
        if (i == np) X_(ub1, i + 1, j, k) += fh * (ind_pos(uh) * xc + ind_neg(uh) * S_(xb, i + 1, j, k));
    }
    {
        const R fl = (j > 1 ? sbar_of(g, a, rhr, h, i, j - 1, k) : R(0)) - sc;
        const R fh = sc - (j < mp ? sbar_of(g, a, rhr, h, i, j + 1, k) : R(0));
        const R ul = Y_(u2, i, j, k), uh = Y_(u2, i, j + 1, k);
        xb_c += fl * (ul < R(0) ? ul : R(0)) + fh * (uh > R(0) ? uh : R(0));
        Y_(ub2, i, j, k) += fl * (ind_pos(ul) * S_(xb, i, j - 1, k) + ind_neg(ul) * xc);
        if (j == mp) Y_(ub2, i, j + 1, k) += fh * (ind_pos(uh) * xc + ind_neg(uh) * S_(xb, i, j + 1, k));
    }
    {
        // z faces 2..lp; the ground cell counts face 2 twice, the sky cell face lp
        if (!gnd) {
            const R c1 = k - 1 == 1 ? R(2) : R(1), c2 = sky ? R(2) : R(1);
            const R fl = c1 * sbar_of(g, a, rhr, h, i, j, k - 1) - c2 * sc;
            const R ul = Z_(u3, i, j, k);
            xb_c += fl * (ul < R(0) ? ul : R(0));
            Z_(ub3, i, j, k) += fl * (ind_pos(ul) * S_(xb, i, j, k - 1) + ind_neg(ul) * xc);
        }
        if (!sky) {
            const R c1 = gnd ? R(2) : R(1), c2 = k + 1 == lp ? R(2) : R(1);
            const R fh = c1 * sc - c2 * sbar_of(g, a, rhr, h, i, j, k + 1);
            const R uh = Z_(u3, i, j, k + 1);
            xb_c += fh * (uh > R(0) ? uh : R(0));
        }
    }
    S_(xbar, i, j, k) = xb_c;
}

#undef CELL_HEAD
#undef S_
#undef X_
#undef Y_
#undef Z_
#undef SP
#undef XP
#undef YP
#undef ZP

inline dim3 cblock() { return dim3(64, 4, 1); }
inline dim3 cgrid(const Geo& g) { return dim3((g.np + 63) / 64, (g.mp + 3) / 4, g.lp); }

int check_adj(mpdata_ctx* c) {
    if (!c) return fail(MPDATA_EINVAL, "null context");
    if (c->real_bytes != 8) return fail(MPDATA_EINVAL, "the adjoint step needs an FP64 context");
    if (c->g.px || c->g.py || c->g.pz)
        return fail(MPDATA_EUNSUPPORTED, "the adjoint step covers open boundaries only");
    if (c->g.ih != 1) return fail(MPDATA_EUNSUPPORTED, "the adjoint step needs ih = 1");
    int dev = -1;
    if (hipGetDevice(&dev) == hipSuccess && dev != c->device)
        return fail(MPDATA_EINVAL, "context was created on device %d, current device is %d",
                    c->device, dev);
    return MPDATA_OK;
}

int ensure_slot(mpdata_ctx* c, TapeSlot& t) {
    int rc;
    if ((rc = ensure(c, &t.xb, SCAL, 8)) || (rc = ensure(c, &t.x0, SCAL, 8)) ||
        (rc = ensure(c, &t.psi, SCAL, 8)) || (rc = ensure(c, &t.cp, SCAL, 8)) ||
        (rc = ensure(c, &t.cn, SCAL, 8)))
        return rc;
    for (int d = 0; d < 3; d++)
        if ((rc = ensure(c, &t.v[d], FACEX + d, 8)) || (rc = ensure(c, &t.fns[d], FACEX + d, 8)) ||
            (rc = ensure(c, &t.vl[d], FACEX + d, 8)))
            return rc;
    return MPDATA_OK;
}

}  // namespace

extern "C" {

int mpdata_adj_reserve(mpdata_ctx* c, int nslots) {
    int rc = check_adj(c);
    if (rc) return rc;
    if (nslots < 1 || nslots > 4096) return fail(MPDATA_EINVAL, "nslots %d outside 1..4096", nslots);
    if ((int)c->tape.size() < nslots) c->tape.resize(nslots);
    for (int q = 0; q < nslots; q++)
        if ((rc = ensure_slot(c, c->tape[q]))) return rc;
    return MPDATA_OK;
}

int mpdata_adj_record_f64(mpdata_ctx* c, int slot, const double* x_in, double* x_out,
                          const double* u1, const double* u2, const double* u3,
                          const double* rhr, const double* h, const double* bcx,
                          const double* bcy, void* stream) {
    int rc = check_adj(c);
    if (rc) return rc;
    if (slot < 0 || slot >= (int)c->tape.size())
        return fail(MPDATA_EINVAL, "tape slot %d outside 0..%d (mpdata_adj_reserve)", slot,
                    (int)c->tape.size() - 1);
    if (!x_in || !x_out || !u1 || !u2 || !u3 || !rhr || !h || !bcx || !bcy)
        return fail(MPDATA_EINVAL, "null pointer");
    const Geo& g = c->g;
    const size_t sb = (size_t)len_of<SCAL>(g) * sizeof(R);
    if ((const char*)x_in < (const char*)x_out + sb && (const char*)x_out < (const char*)x_in + sb)
        return fail(MPDATA_EINVAL, "x_in and x_out overlap");
    Scratch& S = c->prim;
    for (int d = 0; d < 3; d++)
        if ((rc = ensure(c, &S.f[d], FACEX + d, 8))) return rc;
    TapeSlot& t = c->tape[slot];
    hipStream_t s = (hipStream_t)stream;
    using St = Stages<R>;
    R *xb = (R*)t.xb, *x0 = (R*)t.x0, *psi = (R*)t.psi, *cp = (R*)t.cp, *cn = (R*)t.cn;
    R *v[3], *fns[3], *vl[3], *f[3];
    for (int d = 0; d < 3; d++) {
        v[d] = (R*)t.v[d];
        fns[d] = (R*)t.fns[d];
        vl[d] = (R*)t.vl[d];
        f[d] = (R*)S.f[d];
    }
    // the stage sequence of the standard non-oscillatory step
    HIPCHK(hipMemcpyAsync(xb, x_in, sb, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(x0, x_in, sb, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(x_out, x_in, sb, hipMemcpyDeviceToDevice, s));
    HIPCHK(St::putbcinxh(g, xb, bcx, bcy, s));
    HIPCHK(St::upwind(g, u1, u2, u3, xb, psi, rhr, h, s));
    HIPCHK(St::halo_xy(g, psi, St::MIRROR, 1, s));
    HIPCHK(St::pseudo_vel(g, MPDATA_STANDARD, u1, u2, u3, psi, h, v[0], v[1], v[2], fns[0],
                          fns[1], fns[2], s));
    HIPCHK(St::face_bc(g, v[0], v[1], v[2], s));
    HIPCHK(St::face_bc(g, fns[0], fns[1], fns[2], s));
    HIPCHK(St::halo_xy(g, x0, St::COPY, 0, s));
    HIPCHK(St::limiter(g, x0, psi, h, fns[0], fns[1], fns[2], cp, cn, s));
    HIPCHK(St::halo_xy(g, cp, St::ZERO, 1, s));
    HIPCHK(St::halo_xy(g, cn, St::ZERO, 1, s));
    HIPCHK(St::std_flux(g, 1, v[0], v[1], v[2], cp, cn, psi, rhr, f[0], f[1], f[2], vl[0], vl[1],
                        vl[2], s));
    HIPCHK(St::face_bc(g, f[0], f[1], f[2], s));
    HIPCHK(St::face_bc(g, vl[0], vl[1], vl[2], s));
    HIPCHK(St::std_update(g, f[0], f[1], f[2], psi, rhr, h, x_out, s));
    return MPDATA_OK;
}

int mpdata_adj_step_f64(mpdata_ctx* c, int slot, const double* u1, const double* u2,
                        const double* u3, const double* rhr, const double* h, double* xbar,
                        double* ubar1, double* ubar2, double* ubar3, void* stream) {
    int rc = check_adj(c);
    if (rc) return rc;
    if (slot < 0 || slot >= (int)c->tape.size() || !c->tape[slot].xb)
        return fail(MPDATA_EINVAL, "tape slot %d is not reserved", slot);
    if (!u1 || !u2 || !u3 || !rhr || !h || !xbar || !ubar1 || !ubar2 || !ubar3)
        return fail(MPDATA_EINVAL, "null pointer");
    const Geo& g = c->g;
    for (int q = 0; q < A_N; q++)
        if ((rc = ensure(c, &c->adj[q], A_KIND[q], 8))) return rc;
    Adj a{(R*)c->adj[A_PSI], (R*)c->adj[A_X0], (R*)c->adj[A_CP], (R*)c->adj[A_CN],
          (R*)c->adj[A_SUM], (R*)c->adj[A_V1], (R*)c->adj[A_V2], (R*)c->adj[A_V3],
          (R*)c->adj[A_F1], (R*)c->adj[A_F2], (R*)c->adj[A_F3]};
    const TapeSlot& t = c->tape[slot];
    hipStream_t s = (hipStream_t)stream;
    for (int q = 0; q < A_N; q++) {
        const long long n = field_len(g, A_KIND[q]);
        k_zero<<<(int)std::min<long long>(1024, (n + 255) / 256), 256, 0, s>>>((R*)c->adj[q], n);
    }
    HIPCHK(hipGetLastError());
    k_adj_update<<<cgrid(g), cblock(), 0, s>>>(g, xbar, rhr, h, a);
    k_adj_flux<<<cgrid(g), cblock(), 0, s>>>(g, (const R*)t.v[0], (const R*)t.v[1],
                                             (const R*)t.v[2], (const R*)t.vl[0],
                                             (const R*)t.vl[1], (const R*)t.vl[2],
                                             (const R*)t.cp, (const R*)t.cn, (const R*)t.psi, rhr, a);
    k_adj_limiter<<<cgrid(g), cblock(), 0, s>>>(g, (const R*)t.x0, (const R*)t.psi, h,
                                                (const R*)t.fns[0], (const R*)t.fns[1],
                                                (const R*)t.fns[2], (const R*)t.cp,
                                                (const R*)t.cn, a);
    k_adj_pseudo<<<cgrid(g), cblock(), 0, s>>>(g, u1, u2, u3, (const R*)t.psi, h,
                                               (const R*)t.v[0], (const R*)t.v[1],
                                               (const R*)t.v[2], ubar1, ubar2, ubar3, a);
    const int wall_y = g.mp + 2 * g.ih, wall_x = g.np + 2 * g.ih, planes = g.lp + 2 * g.ih;
    k_adj_halo<<<dim3((wall_y + 63) / 64, planes), 64, 0, s>>>(g, 0, a);
    k_adj_halo<<<dim3((wall_x + 63) / 64, planes), 64, 0, s>>>(g, 1, a);
    k_adj_upwind<<<cgrid(g), cblock(), 0, s>>>(g, u1, u2, u3, (const R*)t.xb, rhr, h, xbar, ubar1,
                                               ubar2, ubar3, a);
    HIPCHK(hipGetLastError());
    return MPDATA_OK;
}

}  // extern "C"
