// The fused MPDATA step of the standard antidiffusive variant: donor-cell
// upwind, the standard pseudo-velocities, the non-oscillatory limiter on the
// donor fluxes of the pseudo-velocities, the sign-aware limited velocities
// and the flux-form update of the reference's two-step GPUBC path
// (advection/src_algorithms/upwind3d-gpubc.f90:101-198,
// antidiff3d-standard-gpubc.f90:72-459) in one kernel that reads x, u1, u2,
// u3, h, rhr once and writes the new x once.  The limiter can be off
// (NONOS = false: the pass uses the pseudo-velocities as they are) and any
// direction can be periodic (PER: the clean period-n wrap of the per-stage
// kernels).  Built exactly rounded (mpdata_fused_std.hip: -ffp-contract=off,
// IEEE division; FAST = false) the results are bit-identical to the per-stage
// kernels; the fast build (mpdata_fused_fast.hip: FMA contraction, FAST =
// true, every division a product with a Newton-refined reciprocal) stays
// within 1e-12 of it.
//
// The k-march, work items, tile roles and register queues are those of the
// gauge kernel (mpdata_fused_impl.hpp).  Per plane K the workgroup
//   A  computes xant on plane K over the tile plus a 2-cell ring, and
//      xant/rhr (the ratio every standard flux donates),
//   F  the limited x/y face fluxes of plane Q=K-2,
//   B  the pseudo-velocities V1,V2 on plane P=K-1 and V3 on face K, with the
//      donor fluxes of the pseudo-velocities that the limiter bounds,
//   C  the limiter ratios cp,cn on plane P,
//   D  the new x on plane Q.
// A position's own V1, V2 and V3 stay in registers (the flux through its
// left, lower and bottom faces is formed by the position itself); the
// limiter's donor flux through the right face comes from the lane to the
// right, the one through the upper y-face through LDS.  The inputs of V1
// from column i-1 come from the lane to the left, those of V2 from column
// j-1 through LDS.
#pragma once

#include "mpdata_fused_impl.hpp"

namespace mpdata {

namespace {

// The value held by lane l+1 of the wave (DPP wave_shl:1; lane 63 gets 0).
// Convergent: called with every lane of the wave active.
__device__ __forceinline__ double from_right_lane(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, 0x130, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), 0x130, 0xf, 0xf, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ float from_right_lane(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}

// |v| in the precision of v (the generic lambda below would otherwise see
// only the double overload of fabs at its definition)
__device__ __forceinline__ double abs_(double v) { return __builtin_fabs(v); }
__device__ __forceinline__ float abs_(float v) { return __builtin_fabsf(v); }

// vdyfa and vcorra (mpdata_common.hpp) with the build's division: the same
// operations in the same order when FAST is false
template <bool FAST, typename T>
__device__ __forceinline__ T vdyfa_(T x1, T x2, T a, T rinv) {
    return div_<FAST>((fabs(a) - a * a * rinv) * (fabs(x2) - fabs(x1)),
                      fabs(x2) + fabs(x1) + Eps<T>::v);
}
template <bool FAST, typename T>
__device__ __forceinline__ T vcorra_(T a, T b, T y1, T y2, T rinv) {
    return -div_<FAST>(T(0.125) * a * b * y1 * rinv, y2);
}

// LDS of one workgroup of the standard kernel.
template <typename T, int RX, int RY>
struct StdLds {
    T sx[2][RY + 2][RX + 2];  // x on planes K and K+1 with a 3-cell ring
    T sxa[3][RY][RX];         // xant, planes K-2 .. K
    T src[3][RY][RX];         // xant/rhr, planes K-2 .. K
    T sf1[RY][RX];            // flux through the left x-face, plane Q
    T sf2[RY][RX];            // flux through the lower y-face, plane Q
    T sw2[RY][RX];            // donor flux of V2 through the lower y-face, plane P
    T scp[2][RY][RX];         // cp, cn: planes K-2 and K-1
    T scn[2][RY][RX];
    // h, u1 (face i), u3 (faces P, P+1) of each source column on plane P,
    // read by the position above for V2
    T e_h[RY][RX], e_u1[RY][RX], e_u3a[RY][RX], e_u3b[RY][RX];
};
static_assert(sizeof(StdLds<double, 32, 32>) <= 160 * 1024, "LDS of one CU");

// The standard variant's limited velocity (antidiff3d-standard-gpubc.f90:
// 372-383): pp(v)*(a1*pp(SIGN(1,xm)) + a2*pp(SIGN(1,-xm))) - pn(v)*(a2*pp(
// SIGN(1,xc)) + a1*pp(SIGN(1,-xc))).  One of each pair of pp(SIGN) factors is
// 1 and the other 0, a1, a2 are in [0,1] (never -0), so each bracket is
// a*1 + (+0) = a exactly: the factor of the sign of xm (xc) selects a1 or a2.
template <typename T>
__device__ __forceinline__ T slim(T v, T a1, T a2, T xm, T xc) {
    return hw::pp(v) * (signbit(xm) ? a2 : a1) - hw::pn(v) * (signbit(xc) ? a1 : a2);
}

// One work item (see fused_item in mpdata_fused_impl.hpp for the k-march).
// PER: any of g.px, g.py, g.pz may be set (otherwise they are taken as 0).
// TAPE (limiter on, open boundaries, ih = 1): the step also writes the
// intermediates the adjoint step reads to tp, as the per-stage recording
// leaves them: x with inflow values / edge copies on the wall halo, xant
// with its mirrored x/y halo ring, V on every face of the domain (zero on
// the x/y walls, mirrored on the lids), cp and cn.  Each cell and face is
// written by the one position that owns it.
template <typename T, int RX, int RY, bool NONOS, bool PER, bool TAPE = false, bool FAST = false>
__device__ __forceinline__ void fused_std_item(const Geo& g, const Off& o, int ntx, int tile,
                                               int k0, int k1, const T* __restrict__ x,
                                               const T* __restrict__ u1, const T* __restrict__ u2,
                                               const T* __restrict__ u3,
                                               const T* __restrict__ rhr,
                                               const T* __restrict__ h,
                                               const T* __restrict__ bcx,
                                               const T* __restrict__ bcy, T* __restrict__ xo,
                                               StdLds<T, RX, RY>& L,
                                               const StdTape<T>* tp = nullptr) {
    static_assert(!TAPE || (NONOS && !PER), "the tape covers the limited step with open walls");
    static_assert(!(TAPE && FAST), "the tape is recorded by the exactly rounded build");
    constexpr int NT = RX * RY, TX = RX - 4, TY = RY - 4;
    constexpr int SW = RX + 2, SH = RY + 2, NRING = 2 * SW + 2 * (SH - 2);
    static_assert((64 % RX == 0 || RX % 64 == 0) && NT % 64 == 0, "row layout of the waves");
    static_assert(NRING <= NT, "x ring must fit one extra slot per thread");
    auto& sx = L.sx;
    auto& sxa = L.sxa;
    auto& src = L.src;
    const int np = g.np, mp = g.mp, lp = g.lp;
    const bool px = PER && g.px, py = PER && g.py, pz = PER && g.pz;
    const int t = threadIdx.x;
    const int pi = t % RX, pj = t / RX;  // region index
    const int li = pi - 2, lj = pj - 2;  // tile-local position
    const T ep = Eps<T>::v;
    using hw::donor; using hw::max2; using hw::min2; using hw::min3; using hw::pp; using hw::pn;
    const int SXY = o.sxy, FXY = o.fxy, YXY = o.yxy;

    int rsi = 0, rsj = 0;
    if (t < SW) { rsi = t; rsj = 0; }
    else if (t < 2 * SW) { rsi = t - SW; rsj = SH - 1; }
    else if (t < NRING) { const int r = t - 2 * SW; rsj = 1 + r / 2; rsi = (r & 1) ? SW - 1 : 0; }

    // planes of cells and of z-faces: clamped into the domain on open lids,
    // wrapped with a periodic z (face lp+1 is face 1)
    auto kc = [&](int k) { return pz ? wrap(k, lp) : (k < 1 ? 1 : (k > lp ? lp : k)); };
    auto fc = [&](int k) { return pz ? wrap(k, lp) : (k < 1 ? 1 : (k > lp + 1 ? lp + 1 : k)); };
    // where the x value of region cell (gi, gj) comes from: x (wrapped in
    // periodic directions), the inflow arrays on open walls, else nothing
    auto xsrc = [&](int gi, int gj) -> XSrc<T> {
        if (px) gi = wrap(gi, np);
        if (py) gj = wrap(gj, mp);
        const bool ii = gi >= 1 && gi <= np, jj = gj >= 1 && gj <= mp;
        if (ii && jj) return {x, o.s(gi, gj, 0), SXY};
        if (jj && gi == 0) return {bcx, (gj - 1) - mp, mp};
        if (jj && gi == np + 1) return {bcx, (gj - 1) + mp * lp - mp, mp};
        if (ii && gj == 0) return {bcy, (gi - 1) - np, np};
        if (ii && gj == mp + 1) return {bcy, (gi - 1) + np * lp - np, np};
        return {nullptr, 0, 0};
    };
    auto ld = [&](const XSrc<T>& s, int k) -> T { return s.p ? s.p[s.base + s.stride * kc(k)] : T(0); };

    const int i0 = 1 + (tile % ntx) * TX, j0 = 1 + (tile / ntx) * TY;
    const int gi = i0 + li, gj = j0 + lj;
    const bool in_dom = gi >= 1 && gi <= np && gj >= 1 && gj <= mp;
    // a position in a periodic direction stands for the wrapped cell
    const bool valid = (px || (gi >= 1 && gi <= np)) && (py || (gj >= 1 && gj <= mp));
    const bool do_v1 = valid && (px || gi >= 2) && li >= -1 && lj >= -1 && lj <= TY;
    const bool do_v2 = valid && (py || gj >= 2) && li >= -1 && li <= TX && lj >= -1;
    const bool do_c = valid && li >= -1 && li <= TX && lj >= -1 && lj <= TY;
    const bool do_d = in_dom && li >= 0 && li < TX && lj >= 0 && lj < TY;
    const bool do_f = li >= 0 && li <= TX && lj >= 0 && lj <= TY;
    // open wall faces, where the per-stage path sets the flux to zero
    // (flux_x_bcd / flux_y_bcd): the left x-face / lower y-face of this
    // position is face gi / gj of the domain
    const bool wall1 = !px && (gi <= 1 || gi > np), wall2 = !py && (gj <= 1 || gj > mp);
    // source column: the cell itself, wrapped in periodic directions; on
    // open walls the mirror image on the one-cell ring, clamped beyond
    const int ai = px ? wrap(gi, np) : clampi(mirror(gi, np), 1, np);
    const int aj = py ? wrap(gj, mp) : clampi(mirror(gj, mp), 1, mp);
    // the source column's sx coordinates (a periodic position computes the
    // upwind at its own place, where sx holds the wrapped neighbours)
    const int sai = px ? pi + 1 : ai - i0 + 3, saj = py ? pj + 1 : aj - j0 + 3;
    const XSrc<T> own = valid ? XSrc<T>{x, o.s(ai, aj, 0), SXY} : xsrc(gi, gj);
    const uint32_t so = uint32_t(o.s(ai, aj, 1 - o.ih)) * uint32_t(sizeof(T));
    const uint32_t fo = uint32_t(o.f1(ai, aj, 1 - o.ih)) * uint32_t(sizeof(T));
    // faces i+1 and j+1 of the source column, and x-face i+1 of its column
    // j-1 (wrapped in periodic directions)
    uint32_t fob = 0, sob = 0, fjmb = 0;
    if constexpr (PER) {
        const int ib = (px && ai == np) ? 1 : ai + 1;
        const int jb = (py && aj == mp) ? 1 : aj + 1;
        const int jm = (py && aj == 1) ? mp : aj - 1;
        fob = uint32_t(o.f1(ib, aj, 1 - o.ih)) * uint32_t(sizeof(T));
        sob = uint32_t(o.f2(ai, jb, 1 - o.ih)) * uint32_t(sizeof(T));
        fjmb = uint32_t(o.f1(ib, jm, 1 - o.ih)) * uint32_t(sizeof(T));
    }
    // tape ownership: x/y halo cells of xant and the wall faces n+1 belong to
    // the tile whose output range touches them
    bool own_fx = false, own_fy = false, own_s = false;
    uint32_t tso = 0, tfo = 0;
    if constexpr (TAPE) {
        const bool ci = li >= 0 && li < TX && gi >= 1 && gi <= np;
        const bool cj = lj >= 0 && lj < TY && gj >= 1 && gj <= mp;
        const bool hi = (gi == 0 && li == -1) || (gi == np + 1 && li >= 0 && li <= TX);
        const bool hj = (gj == 0 && lj == -1) || (gj == mp + 1 && lj >= 0 && lj <= TY);
        own_fx = (ci || (gi == np + 1 && li >= 0 && li <= TX)) && cj;
        own_fy = (cj || (gj == mp + 1 && lj >= 0 && lj <= TY)) && ci;
        own_s = (ci || hi) && (cj || hj);
        if (own_s || own_fx || own_fy) {
            tso = uint32_t(o.s(gi, gj, 1 - o.ih)) * uint32_t(sizeof(T));
            tfo = uint32_t(o.f1(gi, gj, 1 - o.ih)) * uint32_t(sizeof(T));
        }
    }
    const int ihm = o.ih - 1;
    auto pl = [&](auto* a, int stride, int k) { return a + (long long)stride * (k + ihm); };
    // loads of the source column's faces i+1 / j+1 and of column j-1 on plane k
    auto ld_u1b = [&](int k) -> T {
        if constexpr (PER) return ldo(pl(u1, FXY, k), fob);
        else return ldo(pl(u1, FXY, k) + 1, fo);
    };
    auto ld_u2b = [&](int k) -> T {
        if constexpr (PER) return ldo(pl(u2, YXY, k), sob);
        else return ldo(pl(u2, YXY, k) + o.yx, so);
    };
    const XSrc<T> ring = t < NRING ? xsrc(i0 + rsi - 3, j0 + rsj - 3) : XSrc<T>{nullptr, 0, 0};
    // x-neighbour availability for the in-plane max/min (edge copies on open
    // walls duplicate the centre and do not change it)
    const bool nxm = px || gi > 1, nxp = px || gi < np, nym = py || gj > 1, nyp = py || gj < mp;

    const int Ks = pz ? k0 - 2 : max(1, k0 - 2), Ke = k1 + 2;
    const int Alast = pz ? k1 + 2 : min(lp, k1 + 2);
    const int Plo = pz ? k0 - 1 : max(1, k0 - 1), Phi = pz ? k1 + 1 : min(lp, k1 + 1);
    // register queues of the source column (as in the gauge kernel) and of
    // the position's own pseudo-velocities: v1q/v2q planes Q and P, v3q
    // faces K-2 and K-1, w3q the donor flux of V3 on the same faces
    T xq[3], u1a[2], u1b[2], u2a[2], u2b[2], u3q[3], hq[2];
    xq[0] = ldo(pl(x, SXY, kc(Ks - 1)), so); xq[1] = ldo(pl(x, SXY, kc(Ks)), so);
    xq[2] = ldo(pl(x, SXY, kc(Ks + 1)), so);
    u1a[0] = ldo(pl(u1, FXY, kc(Ks - 1)), fo); u1a[1] = ldo(pl(u1, FXY, kc(Ks)), fo);
    u1b[0] = ld_u1b(kc(Ks - 1)); u1b[1] = ld_u1b(kc(Ks));
    u2a[0] = ldo(pl(u2, YXY, kc(Ks - 1)), so); u2a[1] = ldo(pl(u2, YXY, kc(Ks)), so);
    u2b[0] = ld_u2b(kc(Ks - 1)); u2b[1] = ld_u2b(kc(Ks));
    u3q[0] = ldo(pl(u3, SXY, fc(Ks - 1)), so); u3q[1] = ldo(pl(u3, SXY, fc(Ks)), so);
    u3q[2] = ldo(pl(u3, SXY, fc(Ks + 1)), so);
    hq[0] = ldo(pl(h, SXY, kc(Ks - 1)), so); hq[1] = ldo(pl(h, SXY, kc(Ks)), so);
    T rq2 = ldo(pl(rhr, SXY, kc(Ks)), so);
    T v1q[2] = {T(0), T(0)}, v2q[2] = {T(0), T(0)};
    T v3q[2] = {T(0), T(0)}, w3q[2] = {T(0), T(0)};
    T xmx[2] = {T(0), T(0)}, xmn[2] = {T(0), T(0)};
    T f3c = T(0);  // flux through the upper z-face of the last output plane

    sx[0][pj + 1][pi + 1] = valid ? xq[1] : ld(own, Ks);
    if (t < NRING) sx[0][rsj][rsi] = ld(ring, Ks);
    __syncthreads();

    uint32_t so_k = so, fo_k = fo;
    auto step = [&](auto phase, int K) __attribute__((always_inline)) {
        constexpr int PH = decltype(phase)::value;
        asm volatile("" : "+s"(K));
        constexpr int Q3a = PH % 3, Q3b = (PH + 1) % 3, Q3c = (PH + 2) % 3;
        constexpr int Q2a = PH % 2, Q2b = (PH + 1) % 2;
        constexpr int R0 = Q3a, R1 = Q3c, R2 = Q3b, E0 = Q2a, E1 = Q2b;
        T& xq1 = xq[Q3a]; T& xq2 = xq[Q3b]; T& xq3 = xq[Q3c];
        T& u1a0 = u1a[Q2a]; T& u1a1 = u1a[Q2b]; T& u1b0 = u1b[Q2a]; T& u1b1 = u1b[Q2b];
        T& u2a0 = u2a[Q2a]; T& u2a1 = u2a[Q2b]; T& u2b0 = u2b[Q2a]; T& u2b1 = u2b[Q2b];
        T& u3q0 = u3q[Q3a]; T& u3q1 = u3q[Q3b]; T& u3q2 = u3q[Q3c];
        T& hq1 = hq[Q2a]; T& hq2 = hq[Q2b];
        T& v1Q = v1q[Q2a]; T& v1P = v1q[Q2b]; T& v2Q = v2q[Q2a]; T& v2P = v2q[Q2b];
        T& v3q1 = v3q[Q2a]; T& v3q2 = v3q[Q2b]; T& w3q1 = w3q[Q2a]; T& w3q2 = w3q[Q2b];
        T& xmx0 = xmx[Q2a]; T& xmn0 = xmn[Q2a]; T& xmx1 = xmx[Q2b]; T& xmn1 = xmn[Q2b];
        asm volatile("" : "+v"(so_k), "+v"(fo_k));
        const uint32_t so = so_k, fo = fo_k;
        const bool A_on = K <= Alast, An_on = K + 1 <= Alast;
        // h and rhr of the output plane Q=K-2 for D
        const T hQ = ldo(pl(h, SXY, kc(K - 2)), so);
        const T rQ = ldo(pl(rhr, SXY, kc(K - 2)), so);
        const T hP = hq1;
        T xr, xown;
        if (An_on) {
            if (t < NRING) xr = ld(ring, K + 1);
            if (!valid) xown = ld(own, K + 1);
        }
        const int P = K - 1;
        const bool BC_on = P >= Plo && P <= Phi;
        // the source column's inputs of V2 at the position above (its
        // column j-1)
        if (BC_on) {
            L.e_h[pj][pi] = hq1;
            L.e_u1[pj][pi] = u1a0;
            L.e_u3a[pj][pi] = u3q0;
            L.e_u3b[pj][pi] = u3q1;
        }

        // ---- A: xant and xant/rhr on plane K ------------------------------
        T xa_new;
        auto stage_a = [&](auto edge) __attribute__((always_inline)) {
            constexpr bool EDGE = decltype(edge)::value;
            auto& sx = L.sx[E0];
            const bool gnd = EDGE && !pz && K == 1, sky = EDGE && !pz && K == lp;
            const T xc = xq2;
            const T hinv = rcp_<FAST>(hq2);
            const T f1p = donor(xc, sx[saj][sai + 1], u1b1);
            const T f1 = donor(sx[saj][sai - 1], xc, u1a1);
            const T f2p = donor(xc, sx[saj + 1][sai], u2b1);
            const T f2 = donor(sx[saj - 1][sai], xc, u2a1);
            T f3p, f3;
            if (gnd) {
                f3p = donor(xc, xq3, u3q2);
                f3 = -f3p;
            } else if (sky) {
                f3 = donor(xq1, xc, u3q1);
                f3p = -f3;
            } else {
                f3p = donor(xc, xq3, u3q2);
                f3 = donor(xq1, xc, u3q1);
            }
            xa_new = rq2 * (xc - (f1p - f1 + f2p - f2 + f3p - f3) * hinv);
            sxa[R0][pj][pi] = xa_new;
            src[R0][pj][pi] = div_<FAST>(xa_new, rq2);
            if constexpr (NONOS) {
                T m = xc, n = xc;
                if (nxm) { T v = sx[pj + 1][pi]; m = max2(m, v); n = min2(n, v); }
                if (nxp) { T v = sx[pj + 1][pi + 2]; m = max2(m, v); n = min2(n, v); }
                if (nym) { T v = sx[pj][pi + 1]; m = max2(m, v); n = min2(n, v); }
                if (nyp) { T v = sx[pj + 2][pi + 1]; m = max2(m, v); n = min2(n, v); }
                if (!gnd) { m = max2(m, xq1); n = min2(n, xq1); }
                if (!sky) { m = max2(m, xq3); n = min2(n, xq3); }
                xmx1 = m;
                xmn1 = n;
            }
        };
        if (A_on) {
            if (!pz && (K == 1 || K == lp)) stage_a(std::true_type{});
            else stage_a(std::false_type{});
        }
        if (An_on) {
            sx[E1][pj + 1][pi + 1] = valid ? xq3 : xown;
            if (t < NRING) sx[E1][rsj][rsi] = xr;
        }
        __syncthreads();
// xant, |xant|, xant/rhr, cp, cn of plane K-d at tile position (a, b)
#define XA(d, a, b) sxa[(d) == 0 ? R0 : ((d) == 1 ? R1 : R2)][(b) + 2][(a) + 2]
#define AX(d, a, b) abs_(XA(d, a, b))
#define RC(d, a, b) src[(d) == 0 ? R0 : ((d) == 1 ? R1 : R2)][(b) + 2][(a) + 2]
#define CP(d, a, b) L.scp[((d) & 1) ? E1 : E0][(b) + 2][(a) + 2]
#define CN(d, a, b) L.scn[((d) & 1) ? E1 : E0][(b) + 2][(a) + 2]
        const int Q = K - 2;
        const bool Dq = Q >= k0 && Q <= k1;
        // ---- limited fluxes through the left x-face and lower y-face of this
        //      position on plane Q (antidiff3d-standard-gpubc.f90:372-407):
        //      V of Q from the previous step, cp/cn of Q and Q-1; formed
        //      before B so that V of Q is dead before V of P is formed ------
        if (Dq && do_f) {
            T F1 = T(0), F2 = T(0);
            if (!wall1) {
                T vl = v1Q;
                if constexpr (NONOS)
                    vl = slim(vl, min3(T(1), CP(2, li, lj), CN(2, li - 1, lj)),
                              min3(T(1), CP(2, li - 1, lj), CN(2, li, lj)), XA(2, li - 1, lj),
                              XA(2, li, lj));
                F1 = donor(RC(2, li - 1, lj), RC(2, li, lj), vl);
            }
            if (!wall2) {
                T vl = v2Q;
                if constexpr (NONOS)
                    vl = slim(vl, min3(T(1), CP(2, li, lj), CN(2, li, lj - 1)),
                              min3(T(1), CP(2, li, lj - 1), CN(2, li, lj)), XA(2, li, lj - 1),
                              XA(2, li, lj));
                F2 = donor(RC(2, li, lj - 1), RC(2, li, lj), vl);
            }
            L.sf1[pj][pi] = F1;
            L.sf2[pj][pi] = F2;
        }
        if constexpr (TAPE)
            if (Dq && own_s) sto(pl(tp->psi, SXY, Q), tso, XA(2, li, lj));
        // ---- B: V3 on face K, V1/V2 on plane P (antidiff3d-standard-
        //      gpubc.f90:72-202), their donor fluxes for the limiter -------
        T v3n = T(0), w3n = T(0);
        if (A_on && do_c && (pz || K >= 2) && K >= k0 - 1) {
            const T hmz = rcp_<FAST>(T(0.5) * (hq1 + hq2));
            const T s1 = u1a0 + u1a1 + u1b1 + u1b0;
            const T s2 = u2a0 + u2b0 + u2b1 + u2a1;
            const T a = u3q1;
            v3n = vdyfa_<FAST>(XA(1, li, lj), xa_new, a, hmz) +
                  vcorra_<FAST>(a, s1,
                         (AX(1, li + 1, lj) + AX(0, li + 1, lj)) - (AX(1, li - 1, lj) + AX(0, li - 1, lj)),
                         AX(1, li + 1, lj) + AX(0, li + 1, lj) + AX(1, li - 1, lj) + AX(0, li - 1, lj) + ep,
                         hmz) +
                  vcorra_<FAST>(a, s2,
                         (AX(1, li, lj + 1) + AX(0, li, lj + 1)) - (AX(1, li, lj - 1) + AX(0, li, lj - 1)),
                         AX(1, li, lj + 1) + AX(0, li, lj + 1) + AX(1, li, lj - 1) + AX(0, li, lj - 1) + ep,
                         hmz);
            if constexpr (NONOS) w3n = donor(XA(1, li, lj), xa_new, v3n);
        }
        // column i-1 of V1 on plane P: the source column of the lane to the
        // left wherever V1 is formed (gi >= 2, or a periodic x)
        const T nv1_h = from_left_lane(hq1);
        const T nv1_u2a = from_left_lane(u2a0);
        const T nv1_u2b = from_left_lane(u2b0);
        const T nv1_u3a = from_left_lane(u3q0);
        const T nv1_u3b = from_left_lane(u3q1);
        T w1 = T(0), w2 = T(0);  // donor fluxes of V1/V2 (left, lower faces of P)
        if (BC_on) {
            const bool zterms = pz || !(P == 1 || P == lp);
            T val1 = T(0), val2 = T(0);
            if (do_v1) {
                const T hmx = rcp_<FAST>(T(0.5) * (nv1_h + hq1));
                const T s2 = nv1_u2a + nv1_u2b + u2b0 + u2a0;
                const T a = u1a0;
                val1 = vdyfa_<FAST>(XA(1, li - 1, lj), XA(1, li, lj), a, hmx) +
                       vcorra_<FAST>(a, s2,
                              (AX(1, li - 1, lj + 1) + AX(1, li, lj + 1)) -
                                  (AX(1, li - 1, lj - 1) + AX(1, li, lj - 1)),
                              AX(1, li - 1, lj + 1) + AX(1, li, lj + 1) + AX(1, li - 1, lj - 1) +
                                  AX(1, li, lj - 1) + ep,
                              hmx);
                if (zterms) {
                    const T s3 = nv1_u3a + nv1_u3b + u3q1 + u3q0;
                    val1 = val1 + vcorra_<FAST>(a, s3,
                                         (AX(0, li - 1, lj) + AX(0, li, lj)) -
                                             (AX(2, li - 1, lj) + AX(2, li, lj)),
                                         AX(0, li - 1, lj) + AX(0, li, lj) + AX(2, li - 1, lj) +
                                             AX(2, li, lj) + ep,
                                         hmx);
                }
                if constexpr (NONOS) w1 = donor(XA(1, li - 1, lj), XA(1, li, lj), val1);
            }
            if (do_v2) {  // pj >= 1 where V2 is formed
                // column j-1: the source column of the position below; u1 on
                // its face i+1 is the next position's face i, except on an
                // open x's last column, whose right neighbour is a mirrored
                // ring position
                const T nv2_h = L.e_h[pj - 1][pi];
                const T nv2_u1a = L.e_u1[pj - 1][pi];
                const T nv2_u3a = L.e_u3a[pj - 1][pi];
                const T nv2_u3b = L.e_u3b[pj - 1][pi];
                T nv2_u1b;
                if (!px && gi == np) {
                    if constexpr (PER) nv2_u1b = ldo(pl(u1, FXY, kc(P)), fjmb);
                    else nv2_u1b = ldo(pl(u1, FXY, kc(P)) - o.fx + 1, fo);
                } else {
                    nv2_u1b = L.e_u1[pj - 1][pi + 1];
                }
                const T hmy = rcp_<FAST>(T(0.5) * (nv2_h + hq1));
                const T s1 = nv2_u1a + u1a0 + u1b0 + nv2_u1b;
                const T a = u2a0;
                val2 = vdyfa_<FAST>(XA(1, li, lj - 1), XA(1, li, lj), a, hmy) +
                       vcorra_<FAST>(a, s1,
                              (AX(1, li + 1, lj - 1) + AX(1, li + 1, lj)) -
                                  (AX(1, li - 1, lj - 1) + AX(1, li - 1, lj)),
                              AX(1, li + 1, lj - 1) + AX(1, li + 1, lj) + AX(1, li - 1, lj - 1) +
                                  AX(1, li - 1, lj) + ep,
                              hmy);
                if (zterms) {
                    const T s3 = nv2_u3a + u3q0 + u3q1 + nv2_u3b;
                    val2 = val2 + vcorra_<FAST>(a, s3,
                                         (AX(0, li, lj - 1) + AX(0, li, lj)) -
                                             (AX(2, li, lj - 1) + AX(2, li, lj)),
                                         AX(0, li, lj - 1) + AX(0, li, lj) + AX(2, li, lj - 1) +
                                             AX(2, li, lj) + ep,
                                         hmy);
                }
                if constexpr (NONOS) w2 = donor(XA(1, li, lj - 1), XA(1, li, lj), val2);
            }
            v1P = val1;
            v2P = val2;
            if constexpr (NONOS) L.sw2[pj][pi] = w2;
            if constexpr (TAPE) {
                if (P >= k0 && P <= k1) {
                    if (own_fx) sto(pl(tp->v[0], FXY, P), tfo, val1);
                    if (own_fy) sto(pl(tp->v[1], YXY, P), tso, val2);
                }
            }
        }
        // the oldest slot of each queue becomes the newest of step K+1
        {
            const int Kn = K + 1;
            xq1 = ldo(pl(x, SXY, kc(Kn + 1)), so);
            u1a0 = ldo(pl(u1, FXY, kc(Kn)), fo);
            u1b0 = ld_u1b(kc(Kn));
            u2a0 = ldo(pl(u2, YXY, kc(Kn)), so);
            u2b0 = ld_u2b(kc(Kn));
            u3q0 = ldo(pl(u3, SXY, fc(Kn + 1)), so);
            hq1 = ldo(pl(h, SXY, kc(Kn)), so);
            rq2 = ldo(pl(rhr, SXY, kc(Kn)), so);
        }
        // donor flux of V1 through the right x-face of P: the left face of
        // the position to the right (the lane to the right, in the same row
        // wherever the limiter runs)
        const T w1p = NONOS ? from_right_lane(w1) : T(0);
        __syncthreads();
        const bool D_on = do_d && Dq;
        // ---- C: cp, cn on plane P (antidiff3d-standard-gpubc.f90:260-347) --
        if constexpr (NONOS) {
            auto stage_c = [&](auto edge) __attribute__((always_inline)) {
                constexpr bool EDGE = decltype(edge)::value;
                const bool gnd = EDGE && !pz && P == 1, sky = EDGE && !pz && P == lp;
                T rp = T(0), rn = T(0);
                if (do_c) {
                    const T xaP = XA(1, li, lj);
                    T mx = XA(1, li - 1, lj), mn = mx;
#define ACC(v, lo, hi) { T t_ = (v); lo = min2(lo, t_); hi = max2(hi, t_); }
                    ACC(xaP, mn, mx) ACC(XA(1, li + 1, lj), mn, mx)
                    ACC(XA(1, li, lj - 1), mn, mx) ACC(XA(1, li, lj + 1), mn, mx)
                    if (!gnd) ACC(XA(2, li, lj), mn, mx)
                    if (!sky) ACC(xa_new, mn, mx)
#undef ACC
                    mx = max2(mx, xmx0);
                    mn = min2(mn, xmn0);
                    const T a1p = w1p, a1 = w1;
                    const T a2p = L.sw2[pj + 1][pi], a2 = w2;
                    T a3p, a3;
                    if (gnd) { a3p = w3n; a3 = -a3p; }
                    else if (sky) { a3 = w3q2; a3p = -a3; }
                    else { a3p = w3n; a3 = w3q2; }
                    rp = div_<FAST>((mx - xaP) * hP,
                                    pn(a1p) + pp(a1) + pn(a2p) + pp(a2) + pn(a3p) + pp(a3) + ep);
                    rn = div_<FAST>((xaP - mn) * hP,
                                    pp(a1p) + pn(a1) + pp(a2p) + pn(a2) + pp(a3p) + pn(a3) + ep);
                }
                L.scp[E1][pj][pi] = rp;
                L.scn[E1][pj][pi] = rn;
            };
            if (BC_on) {
                if (!pz && (P == 1 || P == lp)) stage_c(std::true_type{});
                else stage_c(std::false_type{});
            }
        }
        // limited flux through face Q+1 of the own column (cp/cn of Q from
        // the previous step, of Q+1 from this one)
        auto flux_up = [&]() -> T {
            T vl = v3q2;
            if constexpr (NONOS)
                vl = slim(vl, min3(T(1), CP(1, li, lj), CN(2, li, lj)),
                          min3(T(1), CP(2, li, lj), CN(1, li, lj)), XA(2, li, lj), XA(1, li, lj));
            return donor(RC(2, li, lj), RC(1, li, lj), vl);
        };
        // the lower z-face flux of the first output plane of the range is
        // the upper face of plane k0-1, formed on the step before
        if (Q == k0 - 1 && (pz || Q >= 1) && do_d) f3c = flux_up();
        // ---- D: the new x on plane Q (antidiff3d-standard-gpubc.f90:431-459)
        if (D_on) {
            const bool gnd = !pz && Q == 1, sky = !pz && Q == lp;
            const T F1 = L.sf1[pj][pi], F1p = L.sf1[pj][pi + 1];
            const T F2 = L.sf2[pj][pi], F2p = L.sf2[pj + 1][pi];
            T sum = F1p - F1 + F2p - F2;
            T F3p = T(0);
            if (gnd) {
                F3p = flux_up();
                sum = sum + F3p + F3p;
            } else if (sky) {
                sum = sum - f3c - f3c;
            } else {
                F3p = flux_up();
                sum = sum + F3p - f3c;
            }
            f3c = F3p;
            if constexpr (TAPE) {
                // V3 on the lower face Q of the cell (face 1 = -face 2 on
                // the ground), and on face lp+1 = -face lp on the sky plane;
                // cp and cn of the cell
                T* const tv = pl(tp->v[2], SXY, Q);
                sto(tv, so, gnd ? -v3q2 : v3q1);
                if (sky) sto(tv + SXY, so, -v3q1);
                sto(pl(tp->cp, SXY, Q), so, CP(2, li, lj));
                sto(pl(tp->cn, SXY, Q), so, CN(2, li, lj));
                // x with the inflow values on the wall halo (the upwind's
                // input) and with edge copies there (the limiter's bounds)
                const T xc = ldo(pl(x, SXY, Q), so);
                T* const tb = pl(tp->xb, SXY, Q);
                T* const t0 = pl(tp->x0, SXY, Q);
                sto(tb, so, xc);
                sto(t0, so, xc);
                if (gi == 1) { sto(tb - 1, so, bcx[(gj - 1) + mp * (Q - 1)]); sto(t0 - 1, so, xc); }
                if (gi == np) { sto(tb + 1, so, bcx[(gj - 1) + mp * ((Q - 1) + lp)]); sto(t0 + 1, so, xc); }
                if (gj == 1) { sto(tb - o.sx, so, bcy[(gi - 1) + np * (Q - 1)]); sto(t0 - o.sx, so, xc); }
                if (gj == mp) { sto(tb + o.sx, so, bcy[(gi - 1) + np * ((Q - 1) + lp)]); sto(t0 + o.sx, so, xc); }
            }
            const T hinv = rcp_<FAST>(hQ);
            T* const xoq = pl(xo, SXY, Q);
            sto(xoq, so, rQ * (RC(2, li, lj) - sum * hinv));
            // x halo on open walls: edge copies of x at level n
            const bool ew = !px && (gi == 1 || gi == np), ns = !py && (gj == 1 || gj == mp);
            if (ew || ns) {
                const T xe = ldo(pl(x, SXY, Q), so);
                if (!px && gi == 1) sto(xoq - 1, so, xe);
                if (!px && gi == np) sto(xoq + 1, so, xe);
                if (!py && gj == 1) sto(xoq - o.sx, so, xe);
                if (!py && gj == mp) sto(xoq + o.sx, so, xe);
            }
        }
#undef XA
#undef AX
#undef RC
#undef CP
#undef CN
        v3q1 = v3n;
        w3q1 = w3n;
    };
    for (int K = Ks;;) {
        step(std::integral_constant<int, 0>{}, K);
        if (++K > Ke) break;
        step(std::integral_constant<int, 1>{}, K);
        if (++K > Ke) break;
        step(std::integral_constant<int, 2>{}, K);
        if (++K > Ke) break;
        step(std::integral_constant<int, 3>{}, K);
        if (++K > Ke) break;
        step(std::integral_constant<int, 4>{}, K);
        if (++K > Ke) break;
        step(std::integral_constant<int, 5>{}, K);
        if (++K > Ke) break;
    }
}

template <typename T, int RX, int RY, int MINB, bool NONOS, bool PER, bool TAPE = false,
          bool FAST = false>
__global__ void __launch_bounds__(RX * RY, MINB)
k_fused_std(Geo g, Off o, int ntx, int nty, int nchunk, int clen, const T* __restrict__ x,
            const T* __restrict__ u1, const T* __restrict__ u2, const T* __restrict__ u3,
            const T* __restrict__ rhr, const T* __restrict__ h, const T* __restrict__ bcx,
            const T* __restrict__ bcy, T* __restrict__ xo, StdTape<T> tp) {
    __shared__ StdLds<T, RX, RY> L;
    // work-item order of k_fused_gauge2: contiguous tile runs per XCD
    const int ntile = ntx * nty;
    const int nxcd = gridDim.x >= NXCD ? NXCD : 1;
    const int per_xcd = gridDim.x / nxcd;
    const int xcd = blockIdx.x % nxcd, slot = blockIdx.x / nxcd;
    const int nwg = per_xcd * nxcd;
    if (blockIdx.x >= nwg) return;
    const int tiles_per_xcd = (ntile + nxcd - 1) / nxcd;
    for (int r = slot; ; r += per_xcd) {
        const int chunk = r / tiles_per_xcd, tix = r - chunk * tiles_per_xcd;
        if (chunk >= nchunk) break;
        const int tile = xcd * tiles_per_xcd + tix;
        if (tile >= ntile) continue;
        const int k0 = 1 + chunk * clen, k1 = min(g.lp, k0 + clen - 1);
        fused_std_item<T, RX, RY, NONOS, PER, TAPE, FAST>(g, o, ntx, tile, k0, k1, x, u1, u2, u3,
                                                          rhr, h, bcx, bcy, xo, L, &tp);
    }
}

template <typename T, int RX, int RY, int MINB, bool NONOS, bool PER, bool TAPE = false,
          bool FAST = false>
hipError_t launch_std(const Geo& g, const Off& o, int clen, const T* x, T* xout, const T* u1,
                      const T* u2, const T* u3, const T* rhr, const T* h, const T* bcx,
                      const T* bcy, hipStream_t s, const StdTape<T>& tp = StdTape<T>{}) {
    const int ntx = (g.np + RX - 5) / (RX - 4), nty = (g.mp + RY - 5) / (RY - 4);
    const int nchunk = (g.lp + clen - 1) / clen;
    const int nitem = ntx * nty * nchunk;
    int dev = 0, ncu = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, k_fused_std<T, RX, RY, MINB, NONOS, PER, TAPE, FAST>, RX * RY, 0);
    if (e != hipSuccess) return e;
    const int grid = persistent_grid(nitem, (long long)(ntx * nty + NXCD - 1) / NXCD * nchunk, ncu,
                                     per_cu);
    k_fused_std<T, RX, RY, MINB, NONOS, PER, TAPE, FAST><<<grid, RX * RY, 0, s>>>(
        g, o, ntx, nty, nchunk, clen, x, u1, u2, u3, rhr, h, bcx, bcy, xout, tp);
    return hipGetLastError();
}

// Tiling and k-chunk of the standard kernel: 32x16 tiles (the 32x32 build
// spills registers and measured slower at every size), and the chunk length
// of least modelled wall time (chunk_by_rounds; measured: 128x128x64 78 us
// at 8 planes against 89 at 16; 64x64x32 4 planes; 256x256x128 and
// 512x512x256 64 planes).  MPDATA_FUSED_CFG / _CHUNK override both for
// tests and tuning.
template <typename K>
void choose_std(const Geo& g, K kernel, int& cfg, int& clen) {
    const int env_cfg = fused_config(), env_clen = fused_chunk_env();
    cfg = env_cfg >= 0 ? env_cfg : 1;
    clen = env_clen > 0 ? env_clen : chunk_by_rounds<32, 16>(g, kernel);
    if (clen > g.lp) clen = g.lp;
}

template <typename T, int RX, int RY, int MINB, bool FAST = false>
hipError_t launch_std_any(const Geo& g, const Off& o, int clen, int nonos, const T* x, T* xout,
                          const T* u1, const T* u2, const T* u3, const T* rhr, const T* h,
                          const T* bcx, const T* bcy, hipStream_t s) {
    const bool per = g.px || g.py || g.pz;
    if (nonos) {
        if (per) return launch_std<T, RX, RY, MINB, true, true, false, FAST>(g, o, clen, x, xout, u1, u2, u3, rhr, h, bcx, bcy, s);
        return launch_std<T, RX, RY, MINB, true, false, false, FAST>(g, o, clen, x, xout, u1, u2, u3, rhr, h, bcx, bcy, s);
    }
    if (per) return launch_std<T, RX, RY, MINB, false, true, false, FAST>(g, o, clen, x, xout, u1, u2, u3, rhr, h, bcx, bcy, s);
    return launch_std<T, RX, RY, MINB, false, false, false, FAST>(g, o, clen, x, xout, u1, u2, u3, rhr, h, bcx, bcy, s);
}

// One step of the standard variant on the build FAST selects: the tiling and
// chunk of choose_std, then the halo cells the step leaves to x.
template <typename T, bool FAST>
hipError_t std_step(const Geo& g, int nonos, const T* x, T* xout, const T* u1, const T* u2,
                    const T* u3, const T* rhr, const T* h, const T* bcx, const T* bcy,
                    hipStream_t s) {
    const Off o = make_off(g);
    int cfg = 1, cl = 64;
    choose_std(g, k_fused_std<T, 32, 16, 4, true, false, false, FAST>, cfg, cl);
    hipError_t e =
        cfg == 0
            ? launch_std_any<T, 32, 32, 4, FAST>(g, o, cl, nonos, x, xout, u1, u2, u3, rhr, h, bcx, bcy, s)
            : launch_std_any<T, 32, 16, 4, FAST>(g, o, cl, nonos, x, xout, u1, u2, u3, rhr, h, bcx, bcy, s);
    if (e != hipSuccess) return e;
    k_copy_untouched_halo<T><<<256, 256, 0, s>>>(g, x, xout);
    return hipGetLastError();
}

}  // namespace

}  // namespace mpdata
