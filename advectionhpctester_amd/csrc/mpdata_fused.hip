// Fused single-pass MPDATA step for gfx950: donor-cell upwind, infinite-gauge
// pseudo-velocities, non-oscillatory limiter and the limited corrective update
// of the reference's two-step GPUBC path (advection/src_algorithms/
// upwind3d-gpubc.f90:101-198 and antidiff3d-gauge-gpubc.f90:71-430) in one
// kernel that reads x, u1, u2, u3, h, rhr once and writes the new x once.
//
// Decomposition: one workgroup per TX x TY column tile of the (i,j) plane,
// marching k from the ground to the sky (2.5-D).  Per plane K the workgroup
//   A  computes xant on plane K over the tile plus a 2-cell ring,
//   B  the pseudo-velocities V1,V2 on plane K-1 (tile + 1-face ring) and V3
//      on face K (between planes K-1 and K),
//   C  the limiter ratios cp,cn on plane K-1 (tile + 1-cell ring),
//   D  the new x on plane K-2 (the tile),
// keeping the last three xant/V3/cp/cn planes and two V1/V2 planes in LDS
// rings.  Halo values follow the reference's boundary semantics directly
// (inflow values for the upwind, mirrored xant, zero wall velocities,
// edge-copied x for the limiter, zero cp/cn halo), so no halo-fill kernels
// run.  Expressions are those of the per-stage kernels, operand for operand,
// so the fused result is bit-identical to them.
#include "mpdata_fused.hpp"

namespace mpdata {

namespace {

template <int TX, int TY>
struct Tile {
    static constexpr int NT = TX * TY;
    // LDS regions (local index li = gi - i0; tile cells li in [0,TX))
    static constexpr int AX = TX + 4, AY = TY + 4;   // xant: li in [-2, TX+2)
    static constexpr int V1X = TX + 3, V1Y = TY + 2; // V1 faces li in [-1, TX+2), rows [-1, TY+1)
    static constexpr int V2X = TX + 2, V2Y = TY + 3; // V2 cols [-1, TX+1), faces [-1, TY+2)
    static constexpr int CX = TX + 2, CY = TY + 2;   // V3, cp, cn: [-1, TX+1) x [-1, TY+1)
};

// Offsets with 32-bit arithmetic (the host checks every array fits).
struct Off {
    int sx, sxy;      // scalar strides
    int fx, fxy;      // x-face strides
    int yx, yxy;      // y-face strides
    int ih;
    __device__ int s(int i, int j, int k) const { return (i + ih - 1) + sx * (j + ih - 1) + sxy * (k + ih - 1); }
    __device__ int f1(int i, int j, int k) const { return (i + ih - 1) + fx * (j + ih - 1) + fxy * (k + ih - 1); }
    __device__ int f2(int i, int j, int k) const { return (i + ih - 1) + yx * (j + ih - 1) + yxy * (k + ih - 1); }
    __device__ int f3(int i, int j, int k) const { return s(i, j, k); }  // z-faces share the scalar x/y strides
};

__device__ __forceinline__ int mirror(int i, int n) { return i == 0 ? 2 : (i == n + 1 ? n - 1 : i); }
__device__ __forceinline__ int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }

template <typename T, int TX, int TY>
__global__ void __launch_bounds__(TX * TY)
k_fused_gauge(Geo g, Off o, const T* __restrict__ x, const T* __restrict__ u1,
              const T* __restrict__ u2, const T* __restrict__ u3, const T* __restrict__ rhr,
              const T* __restrict__ h, const T* __restrict__ bcx, const T* __restrict__ bcy,
              T* __restrict__ xo) {
    using TL = Tile<TX, TY>;
    __shared__ T s_xa[3][TL::AY][TL::AX];
    __shared__ T s_v1[2][TL::V1Y][TL::V1X];
    __shared__ T s_v2[2][TL::V2Y][TL::V2X];
    __shared__ T s_v3[3][TL::CY][TL::CX];
    __shared__ T s_cp[3][TL::CY][TL::CX];
    __shared__ T s_cn[3][TL::CY][TL::CX];

    const int np = g.np, mp = g.mp, lp = g.lp;
    const int i0 = 1 + blockIdx.x * TX, j0 = 1 + blockIdx.y * TY;
    const int tid = threadIdx.x;
    const T ep = Eps<T>::v;

    // x as the upwind reads it: inflow values in the open-wall halo
    // (putbcinxh, mpdataoperators.inc:253-287)
    auto x_up = [&](int gi, int gj, int k) -> T {
        if (gi == 0) return bcx[(gj - 1) + mp * ((k - 1))];
        if (gi == np + 1) return bcx[(gj - 1) + mp * ((k - 1) + lp)];
        if (gj == 0) return bcy[(gi - 1) + np * ((k - 1))];
        if (gj == mp + 1) return bcy[(gi - 1) + np * ((k - 1) + lp)];
        return x[o.s(gi, gj, k)];
    };
    // x as the limiter reads it: edge copies (cp_last_to_halo_xy, :551-584)
    auto x_lim = [&](int gi, int gj, int k) -> T {
        return x[o.s(clampi(gi, 1, np), clampi(gj, 1, mp), k)];
    };

    // Halo cells of x_out that the step itself does not define keep their
    // x_in values: the z-halo planes, and the x/y halo lines outside the
    // edge copies written in stage D (the halo corners, and all cells more
    // than one layer out when ih > 1).  Each tile copies the columns of its
    // own tile extended to the domain halo where it touches the domain edge.
    {
        const int ih = g.ih;
        const int ilo = blockIdx.x == 0 ? 1 - ih : i0, ihi = min(i0 + TX - 1, np) + (i0 + TX - 1 >= np ? ih : 0);
        const int jlo = blockIdx.y == 0 ? 1 - ih : j0, jhi = min(j0 + TY - 1, mp) + (j0 + TY - 1 >= mp ? ih : 0);
        const int wi = ihi - ilo + 1, wj = jhi - jlo + 1;
        const int ncol = wi * wj;
        for (int k = 1 - ih; k <= lp + ih; ++k) {
            const bool zhalo = k < 1 || k > lp;
            if (!zhalo && blockIdx.x != 0 && blockIdx.y != 0 && i0 + TX - 1 < np && j0 + TY - 1 < mp)
                continue;  // interior tiles own no x/y halo cells
            for (int q = tid; q < ncol; q += TL::NT) {
                const int gi = ilo + q % wi, gj = jlo + q / wi;
                const bool in_i = gi >= 1 && gi <= np, in_j = gj >= 1 && gj <= mp;
                const bool defined = !zhalo && ((in_i && in_j) || (in_j && (gi == 0 || gi == np + 1)) ||
                                                (in_i && (gj == 0 || gj == mp + 1)));
                if (!defined) xo[o.s(gi, gj, k)] = x[o.s(gi, gj, k)];
            }
        }
    }

    for (int K = 1; K <= lp + 2; ++K) {
        // ---- A: xant on plane K (upwind3d-gpubc.f90:108-198) -----------------
        if (K <= lp) {
            const bool gnd = K == 1, sky = K == lp;
            T(*xa)[TL::AX] = s_xa[K % 3];
            for (int q = tid; q < TL::AX * TL::AY; q += TL::NT) {
                const int lj = q / TL::AX - 2, li = q % TL::AX - 2;
                // out-of-domain positions take the mirrored value
                // (cp_scnd_last_to_halo_xy_fulld, :479-513); positions beyond
                // the one-cell halo are never used and get a clamped copy
                const int gi = clampi(mirror(i0 + li, np), 1, np);
                const int gj = clampi(mirror(j0 + lj, mp), 1, mp);
                const int c = o.s(gi, gj, K);
                T hinv = T(1) / h[c];
                T xc = x[c];
                T f1p = donor(xc, x_up(gi + 1, gj, K), u1[o.f1(gi + 1, gj, K)]);
                T f1 = donor(x_up(gi - 1, gj, K), xc, u1[o.f1(gi, gj, K)]);
                T f2p = donor(xc, x_up(gi, gj + 1, K), u2[o.f2(gi, gj + 1, K)]);
                T f2 = donor(x_up(gi, gj - 1, K), xc, u2[o.f2(gi, gj, K)]);
                T f3p, f3;
                if (gnd) {
                    f3p = donor(xc, x[o.s(gi, gj, K + 1)], u3[o.f3(gi, gj, K + 1)]);
                    f3 = -f3p;
                } else if (sky) {
                    f3 = donor(x[o.s(gi, gj, K - 1)], xc, u3[o.f3(gi, gj, K)]);
                    f3p = -f3;
                } else {
                    f3p = donor(xc, x[o.s(gi, gj, K + 1)], u3[o.f3(gi, gj, K + 1)]);
                    f3 = donor(x[o.s(gi, gj, K - 1)], xc, u3[o.f3(gi, gj, K)]);
                }
                xa[lj + 2][li + 2] = rhr[c] * (xc - (f1p - f1 + f2p - f2 + f3p - f3) * hinv);
            }
        }
        __syncthreads();

        // ---- B: V1,V2 on plane P=K-1, V3 on face F=K (antidiff3d-gauge-gpubc.f90:89-203)
        const int P = K - 1;
#define XA(pl, li, lj) s_xa[(pl) % 3][(lj) + 2][(li) + 2]
        if (P >= 1 && P <= lp) {
            const bool zterms = !(P == 1 || P == lp);
            for (int q = tid; q < TL::V1X * TL::V1Y; q += TL::NT) {
                const int lj = q / TL::V1X - 1, li = q % TL::V1X - 1;  // face li: cells li-1, li
                const int gi = i0 + li, gj = j0 + lj;
                T val = T(0);
                if (gi >= 2 && gi <= np && gj >= 1 && gj <= mp) {
                    T hmx = T(1) / (T(0.5) * (h[o.s(gi - 1, gj, P)] + h[o.s(gi, gj, P)]));
                    T s2 = u2[o.f2(gi - 1, gj, P)] + u2[o.f2(gi - 1, gj + 1, P)] +
                           u2[o.f2(gi, gj + 1, P)] + u2[o.f2(gi, gj, P)];
                    T cr = s2 * rat4(XA(P, li - 1, lj - 1), XA(P, li, lj - 1), XA(P, li - 1, lj + 1),
                                     XA(P, li, lj + 1));
                    if (zterms) {
                        T s3 = u3[o.f3(gi - 1, gj, P)] + u3[o.f3(gi - 1, gj, P + 1)] +
                               u3[o.f3(gi, gj, P + 1)] + u3[o.f3(gi, gj, P)];
                        cr = cr + s3 * rat4(XA(P - 1, li - 1, lj), XA(P - 1, li, lj),
                                            XA(P + 1, li - 1, lj), XA(P + 1, li, lj));
                    }
                    T a = u1[o.f1(gi, gj, P)];
                    val = vdyf(XA(P, li - 1, lj), XA(P, li, lj), a, hmx) - T(0.125) * a * hmx * (cr);
                }
                s_v1[P % 2][lj + 1][li + 1] = val;
            }
            for (int q = tid; q < TL::V2X * TL::V2Y; q += TL::NT) {
                const int lj = q / TL::V2X - 1, li = q % TL::V2X - 1;  // face lj: cells lj-1, lj
                const int gi = i0 + li, gj = j0 + lj;
                T val = T(0);
                if (gj >= 2 && gj <= mp && gi >= 1 && gi <= np) {
                    T hmy = T(1) / (T(0.5) * (h[o.s(gi, gj - 1, P)] + h[o.s(gi, gj, P)]));
                    T s1 = u1[o.f1(gi, gj - 1, P)] + u1[o.f1(gi, gj, P)] + u1[o.f1(gi + 1, gj, P)] +
                           u1[o.f1(gi + 1, gj - 1, P)];
                    T cr = s1 * rat4(XA(P, li - 1, lj - 1), XA(P, li - 1, lj), XA(P, li + 1, lj - 1),
                                     XA(P, li + 1, lj));
                    if (zterms) {
                        T s3 = u3[o.f3(gi, gj - 1, P)] + u3[o.f3(gi, gj, P)] +
                               u3[o.f3(gi, gj, P + 1)] + u3[o.f3(gi, gj - 1, P + 1)];
                        cr = cr + s3 * rat4(XA(P - 1, li, lj - 1), XA(P - 1, li, lj),
                                            XA(P + 1, li, lj - 1), XA(P + 1, li, lj));
                    }
                    T a = u2[o.f2(gi, gj, P)];
                    val = vdyf(XA(P, li, lj - 1), XA(P, li, lj), a, hmy) - T(0.125) * a * hmy * (cr);
                }
                s_v2[P % 2][lj + 1][li + 1] = val;
            }
        }
        if (K >= 2 && K <= lp) {
            const int F = K;  // face between planes F-1 and F
            for (int q = tid; q < TL::CX * TL::CY; q += TL::NT) {
                const int lj = q / TL::CX - 1, li = q % TL::CX - 1;
                const int gi = i0 + li, gj = j0 + lj;
                T val = T(0);
                if (gi >= 1 && gi <= np && gj >= 1 && gj <= mp) {
                    T hmz = T(1) / (T(0.5) * (h[o.s(gi, gj, F - 1)] + h[o.s(gi, gj, F)]));
                    T s1 = u1[o.f1(gi, gj, F - 1)] + u1[o.f1(gi, gj, F)] + u1[o.f1(gi + 1, gj, F)] +
                           u1[o.f1(gi + 1, gj, F - 1)];
                    T s2 = u2[o.f2(gi, gj, F - 1)] + u2[o.f2(gi, gj + 1, F - 1)] +
                           u2[o.f2(gi, gj + 1, F)] + u2[o.f2(gi, gj, F)];
                    T cr = s1 * rat4(XA(F - 1, li - 1, lj), XA(F, li - 1, lj), XA(F - 1, li + 1, lj),
                                     XA(F, li + 1, lj)) +
                           s2 * rat4(XA(F - 1, li, lj - 1), XA(F, li, lj - 1), XA(F - 1, li, lj + 1),
                                     XA(F, li, lj + 1));
                    T a = u3[o.f3(gi, gj, F)];
                    val = vdyf(XA(F - 1, li, lj), XA(F, li, lj), a, hmz) - T(0.125) * a * hmz * (cr);
                }
                s_v3[F % 3][lj + 1][li + 1] = val;
            }
        }
        __syncthreads();

        // ---- C: cp, cn on plane P (antidiff3d-gauge-gpubc.f90:238-324; the
        //      ground plane uses the 6-point minimum of x) --------------------
        if (P >= 1 && P <= lp) {
            const bool gnd = P == 1, sky = P == lp;
            for (int q = tid; q < TL::CX * TL::CY; q += TL::NT) {
                const int lj = q / TL::CX - 1, li = q % TL::CX - 1;
                const int gi = i0 + li, gj = j0 + lj;
                T rp = T(0), rn = T(0);
                if (gi >= 1 && gi <= np && gj >= 1 && gj <= mp) {
                    T mxo = x_lim(gi - 1, gj, P), mno = mxo;
                    T mx = XA(P, li - 1, lj), mn = mx;
#define ACC(v, lo, hi) { T t_ = (v); lo = min2(lo, t_); hi = max2(hi, t_); }
                    ACC(x[o.s(gi, gj, P)], mno, mxo) ACC(x_lim(gi + 1, gj, P), mno, mxo)
                    ACC(x_lim(gi, gj - 1, P), mno, mxo) ACC(x_lim(gi, gj + 1, P), mno, mxo)
                    ACC(XA(P, li, lj), mn, mx) ACC(XA(P, li + 1, lj), mn, mx)
                    ACC(XA(P, li, lj - 1), mn, mx) ACC(XA(P, li, lj + 1), mn, mx)
                    if (!gnd) { ACC(x[o.s(gi, gj, P - 1)], mno, mxo) ACC(XA(P - 1, li, lj), mn, mx) }
                    if (!sky) { ACC(x[o.s(gi, gj, P + 1)], mno, mxo) ACC(XA(P + 1, li, lj), mn, mx) }
#undef ACC
                    mx = max2(mx, mxo);
                    mn = min2(mn, mno);
                    T a1p = s_v1[P % 2][lj + 1][li + 2], a1 = s_v1[P % 2][lj + 1][li + 1];
                    T a2p = s_v2[P % 2][lj + 2][li + 1], a2 = s_v2[P % 2][lj + 1][li + 1];
                    T a3p, a3;
                    if (gnd) {
                        a3p = s_v3[(P + 1) % 3][lj + 1][li + 1];
                        a3 = -a3p;
                    } else if (sky) {
                        a3 = s_v3[P % 3][lj + 1][li + 1];
                        a3p = -a3;
                    } else {
                        a3p = s_v3[(P + 1) % 3][lj + 1][li + 1];
                        a3 = s_v3[P % 3][lj + 1][li + 1];
                    }
                    T xc = XA(P, li, lj), hc = h[o.s(gi, gj, P)];
                    rp = (mx - xc) * hc / (pn(a1p) + pp(a1) + pn(a2p) + pp(a2) + pn(a3p) + pp(a3) + ep);
                    rn = (xc - mn) * hc / (pp(a1p) + pn(a1) + pp(a2p) + pn(a2) + pp(a3p) + pn(a3) + ep);
                }
                s_cp[P % 3][lj + 1][li + 1] = rp;
                s_cn[P % 3][lj + 1][li + 1] = rn;
            }
        }
        __syncthreads();

        // ---- D: limited fluxes and the new x on plane Q=K-2
        //      (antidiff3d-gauge-gpubc.f90:340-428) --------------------------
        const int Q = K - 2;
        if (Q >= 1) {
            const bool gnd = Q == 1, sky = Q == lp;
            const int li = tid % TX, lj = tid / TX;
            const int gi = i0 + li, gj = j0 + lj;
            if (gi <= np && gj <= mp) {
#define CP(pl, a, b) s_cp[(pl) % 3][(b) + 1][(a) + 1]
#define CN(pl, a, b) s_cn[(pl) % 3][(b) + 1][(a) + 1]
                const T v1c = s_v1[Q % 2][lj + 1][li + 1], v1p = s_v1[Q % 2][lj + 1][li + 2];
                const T v2c = s_v2[Q % 2][lj + 1][li + 1], v2p = s_v2[Q % 2][lj + 2][li + 1];
                T F1 = gflux(v1c, CP(Q, li, lj), CN(Q, li - 1, lj), CP(Q, li - 1, lj), CN(Q, li, lj));
                T F1p = gflux(v1p, CP(Q, li + 1, lj), CN(Q, li, lj), CP(Q, li, lj), CN(Q, li + 1, lj));
                T F2 = gflux(v2c, CP(Q, li, lj), CN(Q, li, lj - 1), CP(Q, li, lj - 1), CN(Q, li, lj));
                T F2p = gflux(v2p, CP(Q, li, lj + 1), CN(Q, li, lj), CP(Q, li, lj), CN(Q, li, lj + 1));
                T F3, F3p;
                if (gnd) {
                    const T v = s_v3[(Q + 1) % 3][lj + 1][li + 1];
                    F3 = gflux(-v, CP(Q, li, lj), CN(Q + 1, li, lj), CP(Q + 1, li, lj), CN(Q, li, lj));
                    F3p = gflux(v, CP(Q + 1, li, lj), CN(Q, li, lj), CP(Q, li, lj), CN(Q + 1, li, lj));
                } else if (sky) {
                    const T v = s_v3[Q % 3][lj + 1][li + 1];
                    F3 = gflux(v, CP(Q, li, lj), CN(Q - 1, li, lj), CP(Q - 1, li, lj), CN(Q, li, lj));
                    F3p = gflux(-v, CP(Q - 1, li, lj), CN(Q, li, lj), CP(Q, li, lj), CN(Q - 1, li, lj));
                } else {
                    const T vc = s_v3[Q % 3][lj + 1][li + 1], vp = s_v3[(Q + 1) % 3][lj + 1][li + 1];
                    F3 = gflux(vc, CP(Q, li, lj), CN(Q - 1, li, lj), CP(Q - 1, li, lj), CN(Q, li, lj));
                    F3p = gflux(vp, CP(Q + 1, li, lj), CN(Q, li, lj), CP(Q, li, lj), CN(Q + 1, li, lj));
                }
#undef CP
#undef CN
                const int c = o.s(gi, gj, Q);
                const T r = rhr[c];
                xo[c] = r * (XA(Q, li, lj) / r - (F1p - F1 + F2p - F2 + F3p - F3) / h[c]);
                // x halo as the reference leaves it: edge copies of x^n
                // written before the update (cp_last_to_halo_xy)
                if (gi == 1) xo[o.s(0, gj, Q)] = x[c];
                if (gi == np) xo[o.s(np + 1, gj, Q)] = x[c];
                if (gj == 1) xo[o.s(gi, 0, Q)] = x[c];
                if (gj == mp) xo[o.s(gi, mp + 1, Q)] = x[c];
            }
        }
#undef XA
        __syncthreads();
    }
}

constexpr int FTX = 32, FTY = 8;

}  // namespace

template <typename T>
bool Fused<T>::supports(const Geo& g, int variant, int iters, int nonos) {
    if (variant != 1 || iters != 2 || nonos != 1) return false;
    if (g.px || g.py || g.pz) return false;
    // 32-bit offsets inside the kernel
    const long long biggest = len_of<FACEX>(g) > len_of<FACEY>(g) ? len_of<FACEX>(g) : len_of<FACEY>(g);
    return biggest < (1ll << 31) - 1 && len_of<FACEZ>(g) < (1ll << 31) - 1;
}

template <typename T>
const char* Fused<T>::scope() {
    return "the gauge variant with 2 iterations, the limiter on and open boundaries";
}

template <typename T>
hipError_t Fused<T>::step(const Geo& g, const T* x, T* xout, const T* u1, const T* u2,
                          const T* u3, const T* rhr, const T* h, const T* bcx, const T* bcy,
                          hipStream_t s) {
    Off o;
    o.ih = g.ih;
    o.sx = g.np + 2 * g.ih;
    o.sxy = o.sx * (g.mp + 2 * g.ih);
    o.fx = g.np + 1 + 2 * g.ih;
    o.fxy = o.fx * (g.mp + 2 * g.ih);
    o.yx = g.np + 2 * g.ih;
    o.yxy = o.yx * (g.mp + 1 + 2 * g.ih);
    dim3 grid((g.np + FTX - 1) / FTX, (g.mp + FTY - 1) / FTY, 1);
    k_fused_gauge<T, FTX, FTY><<<grid, FTX * FTY, 0, s>>>(g, o, x, u1, u2, u3, rhr, h, bcx, bcy, xout);
    return hipGetLastError();
}

template struct Fused<double>;
template struct Fused<float>;

}  // namespace mpdata
