// Kernel templates of the fused single-pass MPDATA step, shared by the
// exactly rounded build (mpdata_fused.hip, -ffp-contract=off, IEEE division)
// and the fast one (mpdata_fused_fast.hip, FMA contraction and
// Newton-refined reciprocals, FAST = true).
//
// Fused single-pass MPDATA step for gfx950: donor-cell upwind, infinite-gauge
// pseudo-velocities, non-oscillatory limiter and the limited corrective update
// of the reference's two-step GPUBC path (advection/src_algorithms/
// upwind3d-gpubc.f90:101-198 and antidiff3d-gauge-gpubc.f90:71-430) in one
// kernel that reads x, u1, u2, u3, h, rhr once and writes the new x once.
// The fixed-position kernel also has a density mode (DensIO) for the step of
// mpdata3d-rho-gauge-gpubc.f90:107-622.
//
// Decomposition: one workgroup per TX x TY column tile of the (i,j) plane,
// marching k from the ground to the sky (2.5-D).  Per plane K the workgroup
//   A  computes xant on plane K over the tile plus a 2-cell ring,
//   B  the pseudo-velocities V1,V2 on plane K-1 (tile + 1-face ring) and V3
//      on face K (between planes K-1 and K),
//   C  the limiter ratios cp,cn on plane K-1 (tile + 1-cell ring),
//   D  the new x on plane K-2 (the tile),
// keeping the last three xant planes and two V1/V2/cp/cn planes in LDS
// rings and the V3 faces and the column's inputs in register queues.  Halo
// values follow the reference's boundary semantics directly (inflow values
// for the upwind, mirrored xant, zero wall velocities, edge-copied x for
// the limiter, zero cp/cn halo), so no halo-fill kernels run.  Expressions are those of the per-stage kernels, operand for operand,
// so the fused result is bit-identical to them.
#pragma once
#include "mpdata_fused.hpp"

#include <algorithm>
#include <type_traits>
#include <cstdlib>
#include <cstddef>

namespace mpdata {

namespace {

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

// Element at byte offset `off` from `p`, where p is uniform over the
// workgroup (an array's plane) and off a per-lane 32-bit offset: the
// global_load/store saddr form (scalar base + 32-bit vector offset), with no
// per-lane 64-bit address arithmetic.
template <typename T>
__device__ __forceinline__ T ldo(const T* p, uint32_t off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + off);
}
template <typename T>
__device__ __forceinline__ void sto(T* p, uint32_t off, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + off) = v;
}

// The value held by lane l-1 of the wave (DPP wave_shr:1; lane 0, which has
// no source lane, gets 0 through bound_ctrl, so no old value has to be set).
// Convergent: called with every lane of the wave active.
__device__ __forceinline__ double from_left_lane(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, 0x138, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), 0x138, 0xf, 0xf, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ float from_left_lane(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }

// Reciprocal and quotient of the fast build (FAST = true): the hardware
// reciprocal (v_rcp) refined by one Newton step e = 1 - b*r, r += e*r, and
// the quotient a*r: a few ulps from the IEEE quotient, in four
// instructions instead of the ten of the IEEE expansion with its scale and
// fix-up steps.  The 50-step difference from the exact build at
// 512x512x256 is 1.2e-14 (6.6e-15 with a second Newton step and a
// correction of the quotient, 1.9 % slower), against the 1e-12 tolerance.  The operands are positive and of order one (h, rhr and the
// limiter denominators, which include eps), far from the ends of the
// exponent range.  FAST = false is the reference's IEEE division.
__device__ __forceinline__ double rcp_nr(double b) {
    const double r = __builtin_amdgcn_rcp(b);
    return __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
}
__device__ __forceinline__ float rcp_nr(float b) {
    const float r = __builtin_amdgcn_rcpf(b);
    return __builtin_fmaf(__builtin_fmaf(-b, r, 1.0f), r, r);
}
__device__ __forceinline__ double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <bool FAST, typename T>
__device__ __forceinline__ T rcp_(T b) {
    if constexpr (FAST) return rcp_nr(b);
    else return T(1) / b;
}
template <bool FAST, typename T>
__device__ __forceinline__ T div_(T a, T b) {
    if constexpr (FAST) {
        return a * rcp_nr(b);
    } else {
        return a / b;
    }
}

// ---------------------------------------------------------------------------
// The fused kernel: fixed per-thread column positions.
//
// A workgroup of RX*RY threads covers an RX x RY region of (i,j) columns:
// the TX x TY output tile (TX=RX-4, TY=RY-4) plus a 2-cell ring.  Thread t
// owns region position (t % RX, t / RX) for the whole k-march, so every value
// a position needs only at its own column (x, u, h, rhr of the planes around
// K, its xant, V3, pair sums) lives in a register queue, and only values
// read by in-plane neighbours go through LDS:
//   sx   x on planes K and K+1 with a 3-cell ring (inflow halo values)
//   sxa  xant, planes K-2 .. K
//   sv1, sv2   V1 and V2, planes K-2 and K-1
//   scp, scn   cp and cn, planes K-2 and K-1
//   sf1, sf2   limited x/y face fluxes of plane K-2
//   e_*  (1024-thread tilings) h, u1, u3 of each column on plane K-1
// The z cross terms of V1 and V2 use pair sums formed one plane at a time:
// rat4(xa(i-1,k-1), xa(i,k-1), xa(i-1,k+1), xa(i,k+1)) =
//   ((xa(i,k+1)+xa(i-1,k+1)) - (xa(i,k-1)+xa(i-1,k-1)))*0.25,
// the same operations in the same order as the reference.
//
// Two barriers per plane: stage A reads the x plane that the previous step
// wrote and writes the next one; B reads the xant A wrote; C/D read the V
// that B wrote.  A slot of a ring is rewritten only after a barrier that
// every reader of its previous plane has passed (three xant slots, because
// A of plane K runs while slower threads may still be in C of plane K-1,
// reading xant K-2).
//
// The grid is persistent: workgroups loop over (tile, k-chunk) work items,
// a chunk of output planes [k0,k1] recomputing xant from k0-2 and cp,cn,V
// from k0-1.  Inputs of plane K+1 are loaded while plane K is computed.
// ---------------------------------------------------------------------------
template <typename T>
struct XSrc {  // where one sx cell comes from: x, bcx, bcy or nothing
    const T* p;
    int base, stride;
};

// Whether a tiling passes the j-1 column inputs of V2 through LDS (the
// 1024-thread tilings, one workgroup per CU, have the room; the 512-thread
// ones run two workgroups per CU and load them).
template <int RX, int RY>
struct XJ { static constexpr bool value = RX * RY >= 1024; };

// LDS of one workgroup of the fixed-position kernel.
template <typename T, int RX, int RY>
struct FusedLds {
    static constexpr int NE = XJ<RX, RY>::value ? RY * RX : 1;
    T sx[2][RY + 2][RX + 2];
    T sxa[3][RY][RX];
    T sv1[2][RY][RX];
    T sv2[2][RY][RX];
    T scp[2][RY][RX];
    T scn[2][RY][RX];
    T sf1[RY][RX];  // limited flux through the left x-face, plane K-2
    T sf2[RY][RX];  // limited flux through the lower y-face, plane K-2
    // h, u1 (face i), u3 (faces P, P+1) of each source column on plane P
    T e_h[NE], e_u1[NE], e_u3a[NE], e_u3b[NE];
};
static_assert(sizeof(FusedLds<double, 32, 32>) <= 160 * 1024, "LDS of one CU");

// One work item: output tile `tile` (ntx tiles per row) over the output
// planes [k0, k1].  Every argument is uniform over the workgroup, so all
// threads pass the same barriers.
// DENS: the density step (DensIO) -- the upwind divides by h without the
// rhr factor, the update has no rhr, and the output plane also writes
// x_incomp, pexr1, pexr2, rhoadv, rhr and the total mass fluxes
// (mpdata3d-rho-gauge-gpubc.f90:123-192, 377-593).
// FAST: the fast build's division (rcp_, div_); the file that instantiates it
// is compiled with FMA contraction.
// PASS: 0 the two-pass step; 1 the same, also writing the limited
// velocities of the corrective pass (the fluxes F of the infinite gauge) to
// pio->ul1..3 on faces 1..n+1, as the reference's stage returns them for a
// further pass; 2 a further corrective pass (antidiff3d-gauge-gpubc.f90:81-
// 83): stage A reads the previous pass's result pio->xa (at the mirrored
// source column) instead of computing the upwind, u1..u3 are that pass's
// limited velocities and x is the field at time level n (limiter bounds).
// TAPE (the two-pass step with open walls, ih = 1): the step also writes the
// intermediates the adjoint step reads to tp, as the per-stage recording
// leaves them: x with inflow values / edge copies on the wall halo, xant
// with its mirrored x/y halo ring, V on every face of the domain (zero on
// the x/y walls, mirrored on the lids), cp and cn.  Each cell and face is
// written by the one position that owns it.
template <typename T, int RX, int RY, bool DENS = false, bool FAST = false, int PASS = 0,
          bool TAPE = false>
__device__ __forceinline__ void fused_item(const Geo& g, const Off& o, int ntx, int tile, int k0,
                                           int k1, const T* __restrict__ x, const T* __restrict__ u1,
                                           const T* __restrict__ u2, const T* __restrict__ u3,
                                           const T* __restrict__ rhr, const T* __restrict__ h,
                                           const T* __restrict__ bcx, const T* __restrict__ bcy,
                                           T* __restrict__ xo,
                                           FusedLds<T, RX, RY>& L,
                                           const DensIO<T>* dio = nullptr,
                                           const PassIO<T>* pio = nullptr,
                                           const StdTape<T>* tp = nullptr) {
    static_assert(!TAPE || (!DENS && !FAST && PASS == 0), "the tape covers the two-pass step");
    constexpr int NT = RX * RY, TX = RX - 4, TY = RY - 4;
    constexpr int SW = RX + 2, SH = RY + 2, NRING = 2 * SW + 2 * (SH - 2);
    // lane t-1 holds region position (pi-1, pj) for every pi >= 1, in the
    // same wave, and the workgroup is whole waves
    static_assert((64 % RX == 0 || RX % 64 == 0) && NT % 64 == 0, "row layout of the waves");
    static_assert(NRING <= NT, "x ring must fit one extra slot per thread");
    auto& sx = L.sx;
    auto& sxa = L.sxa;
    auto& sv1 = L.sv1;
    const int np = g.np, mp = g.mp, lp = g.lp;
    const int t = threadIdx.x;
    const int pi = t % RX, pj = t / RX;  // region index
    const int li = pi - 2, lj = pj - 2;  // tile-local position
    const T ep = Eps<T>::v;
    using hw::donor; using hw::gflux; using hw::max2; using hw::min2; using hw::pp; using hw::pn;
    const int SXY = o.sxy, FXY = o.fxy, YXY = o.yxy;

    // ring slot of the sx plane handled by this thread (if t < NRING)
    int rsi = 0, rsj = 0;
    if (t < SW) { rsi = t; rsj = 0; }
    else if (t < 2 * SW) { rsi = t - SW; rsj = SH - 1; }
    else if (t < NRING) { const int r = t - 2 * SW; rsj = 1 + r / 2; rsi = (r & 1) ? SW - 1 : 0; }

    auto xsrc = [&](int gi, int gj) -> XSrc<T> {
        const bool ii = gi >= 1 && gi <= np, jj = gj >= 1 && gj <= mp;
        if (ii && jj) return {x, o.s(gi, gj, 0), SXY};
        if (jj && gi == 0) return {bcx, (gj - 1) - mp, mp};
        if (jj && gi == np + 1) return {bcx, (gj - 1) + mp * lp - mp, mp};
        if (ii && gj == 0) return {bcy, (gi - 1) - np, np};
        if (ii && gj == mp + 1) return {bcy, (gi - 1) + np * lp - np, np};
        return {nullptr, 0, 0};
    };
    auto ld = [](const XSrc<T>& s, int k) -> T { return s.p ? s.p[s.base + s.stride * k] : T(0); };
    auto kc = [&](int k) { return k < 1 ? 1 : (k > lp ? lp : k); };        // cell planes
    auto fc = [&](int k) { return k < 1 ? 1 : (k > lp + 1 ? lp + 1 : k); };  // z-face planes


    const int i0 = 1 + (tile % ntx) * TX, j0 = 1 + (tile / ntx) * TY;
    const int gi = i0 + li, gj = j0 + lj;
    const bool in_dom = gi >= 1 && gi <= np && gj >= 1 && gj <= mp;
    // roles of this position
    const bool do_v1 = in_dom && gi >= 2 && li >= -1 && lj >= -1 && lj <= TY;
    const bool do_v2 = in_dom && gj >= 2 && li >= -1 && li <= TX && lj >= -1;
    const bool do_c = in_dom && li >= -1 && li <= TX && lj >= -1 && lj <= TY;
    const bool do_d = in_dom && li >= 0 && li < TX && lj >= 0 && lj < TY;
    // left/lower face fluxes: every x-face and y-face of a tile cell
    const bool do_f = li >= 0 && li <= TX && lj >= 0 && lj <= TY;
    // upwind source cell: itself in the domain, the mirror image on the
    // one-cell halo ring (cp_scnd_last_to_halo_xy_fulld), clamped beyond
    const int ai = clampi(mirror(gi, np), 1, np), aj = clampi(mirror(gj, mp), 1, mp);
    const int sai = ai - i0 + 3, saj = aj - j0 + 3;  // its sx coordinates
    const int bs = o.s(ai, aj, 0);
    const XSrc<T> own = in_dom ? XSrc<T>{x, bs, SXY} : xsrc(gi, gj);
    // byte offsets of the source column inside a plane: so for cell arrays,
    // z-faces and y-faces (whose in-plane strides equal the cell ones),
    // fo for x-faces; planes are addressed by uniform pointers
    const uint32_t so = uint32_t(o.s(ai, aj, 1 - o.ih)) * uint32_t(sizeof(T));
    const uint32_t fo = uint32_t(o.f1(ai, aj, 1 - o.ih)) * uint32_t(sizeof(T));
    const int ihm = o.ih - 1;
    auto pl = [&](auto* a, int stride, int k) { return a + (long long)stride * (k + ihm); };
    const XSrc<T> ring = t < NRING ? xsrc(i0 + rsi - 3, j0 + rsj - 3) : XSrc<T>{nullptr, 0, 0};
    // x-neighbour availability for the in-plane max/min (edge copies
    // duplicate the centre and do not change it)
    const bool nxm = gi > 1, nxp = gi < np, nym = gj > 1, nyp = gj < mp;
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

    const int Ks = max(1, k0 - 2), Ke = k1 + 2;
    const int Alast = min(lp, k1 + 2);
    // register queues of the source column; during step K they hold
    //   xq1..3 = x(K-1..K+1), u1a0/1 = u1(i,K-1..K), u1b0/1 = u1(i+1,K-1..K),
    //   u2a/u2b likewise at j and j+1, u3q0..2 = faces K-1..K+1,
    //   hq1..2 = h(K-1..K), rq2 = rhr(K), v3q1..2 = V3 faces K-2..K-1,
    //   xmx0/xmn0 = bounds of x for the limiter on plane K-1 (xmx1/xmn1:
    //   plane K, formed in A).
    // The column's own xant of planes K-2..K comes from sxa: the slot of
    // an own position is rewritten only by this thread, later in program
    // order.
    // The k loop is unrolled six-fold: step PH of a round handles a plane
    // K with K - Ks = PH (mod 6), so every ring slot (K - Ks mod 2 or 3) is
    // a compile-time constant, which makes the LDS offsets immediates, and
    // the queues are arrays whose roles rotate with PH instead of being
    // copied: a new plane is loaded into the slot of the oldest one.
    T xq[3], u1a[2], u1b[2], u2a[2], u2b[2], u3q[3], hq[2];
    xq[0] = ldo(pl(x, SXY, kc(Ks - 1)), so); xq[1] = ldo(pl(x, SXY, kc(Ks)), so);
    xq[2] = ldo(pl(x, SXY, kc(Ks + 1)), so);
    u1a[0] = ldo(pl(u1, FXY, kc(Ks - 1)), fo); u1a[1] = ldo(pl(u1, FXY, kc(Ks)), fo);
    u1b[0] = ldo(pl(u1, FXY, kc(Ks - 1)) + 1, fo); u1b[1] = ldo(pl(u1, FXY, kc(Ks)) + 1, fo);
    u2a[0] = ldo(pl(u2, YXY, kc(Ks - 1)), so); u2a[1] = ldo(pl(u2, YXY, kc(Ks)), so);
    u2b[0] = ldo(pl(u2, YXY, kc(Ks - 1)) + o.yx, so); u2b[1] = ldo(pl(u2, YXY, kc(Ks)) + o.yx, so);
    u3q[0] = ldo(pl(u3, SXY, fc(Ks - 1)), so); u3q[1] = ldo(pl(u3, SXY, fc(Ks)), so);
    u3q[2] = ldo(pl(u3, SXY, fc(Ks + 1)), so);
    hq[0] = ldo(pl(h, SXY, kc(Ks - 1)), so); hq[1] = ldo(pl(h, SXY, kc(Ks)), so);
    T rq2 = T(0);
    if constexpr (!DENS && PASS != 2) rq2 = ldo(pl(rhr, SXY, kc(Ks)), so);
    // PASS 2: the previous pass's result on plane K (stage A's xant)
    T xaq = T(0);
    if constexpr (PASS == 2) xaq = ldo(pl(pio->xa, SXY, kc(Ks)), so);
    T v3q[2] = {T(0), T(0)};
    T xmx[2] = {T(0), T(0)}, xmn[2] = {T(0), T(0)};
    T f3c = T(0);  // limited flux through the upper z-face of the last output plane

    // x of the first plane into LDS; the barrier also orders this item
    // after every LDS read of the previous one
    sx[0][pj + 1][pi + 1] = in_dom ? xq[1] : ld(own, Ks);
    if (t < NRING) sx[0][rsj][rsi] = ld(ring, Ks);
    __syncthreads();

    // The tile arrays from sv2 on lie beyond the 64 KiB that a ds
    // instruction's immediate offset reaches from the lane's position.  They
    // are addressed from two per-lane byte offsets made opaque here, so that
    // every access is one of them plus an immediate; otherwise the compiler
    // keeps one address register per array and ring slot through the loop.
    using LT = FusedLds<T, RX, RY>;
    constexpr int SZ = (int)sizeof(T), PLN = RY * RX * SZ;  // one plane of a tile array
    constexpr int MID = (int)offsetof(LT, sv2);
    constexpr int HI = (int)offsetof(LT, e_h) - RX * SZ;  // e arrays are read one row back
    uint32_t lmid = uint32_t(MID + t * SZ), lhi = uint32_t(HI + t * SZ);
    asm volatile("" : "+v"(lmid), "+v"(lhi));
    char* const lds0 = reinterpret_cast<char*>(&L);
    // array `arr` at ring slot s, offset (db rows, da columns) from the lane
#define LMID(arr, s, db, da) \
    (*reinterpret_cast<T*>(lds0 + lmid + ((int)offsetof(LT, arr) - MID + (s) * PLN + ((db) * RX + (da)) * SZ)))
#define LHI(arr, r) (*reinterpret_cast<T*>(lds0 + lhi + ((int)offsetof(LT, arr) - HI + (r) * SZ)))

    uint32_t so_k = so, fo_k = fo;
    auto step = [&](auto phase, int K) __attribute__((always_inline)) {
        constexpr int PH = decltype(phase)::value;
        // K opaque: every plane-dependent quantity is formed within the
        // step, rather than as one induction variable per unrolled step
        // held in registers through the loop
        asm volatile("" : "+s"(K));
        // queue slots: oldest, middle, newest of a three-plane queue; older,
        // newer of a two-plane one
        constexpr int Q3a = PH % 3, Q3b = (PH + 1) % 3, Q3c = (PH + 2) % 3;
        constexpr int Q2a = PH % 2, Q2b = (PH + 1) % 2;
        // LDS ring slots of plane K-d: mod 3 (R0, R1, R2 for d = 0, 1, 2) and
        // mod 2 (E0 for even d, E1 for odd d)
        constexpr int R0 = Q3a, R1 = Q3c, R2 = Q3b, E0 = Q2a, E1 = Q2b;
        T& xq1 = xq[Q3a]; T& xq2 = xq[Q3b]; T& xq3 = xq[Q3c];
        T& u1a0 = u1a[Q2a]; T& u1a1 = u1a[Q2b]; T& u1b0 = u1b[Q2a]; T& u1b1 = u1b[Q2b];
        T& u2a0 = u2a[Q2a]; T& u2a1 = u2a[Q2b]; T& u2b0 = u2b[Q2a]; T& u2b1 = u2b[Q2b];
        T& u3q0 = u3q[Q3a]; T& u3q1 = u3q[Q3b]; T& u3q2 = u3q[Q3c];
        T& hq1 = hq[Q2a]; T& hq2 = hq[Q2b];
        T& v3q1 = v3q[Q2a]; T& v3q2 = v3q[Q2b];
        T& xmx0 = xmx[Q2a]; T& xmn0 = xmn[Q2a]; T& xmx1 = xmx[Q2b]; T& xmn1 = xmn[Q2b];
        // the column offsets are opaque to the optimiser at each plane, so
        // loads keep the saddr form (plane pointer + offset) instead of
        // becoming per-lane 64-bit pointer inductions
        asm volatile("" : "+v"(so_k), "+v"(fo_k));
        const uint32_t so = so_k, fo = fo_k;
        const bool A_on = K <= Alast, An_on = K + 1 <= Alast;
        // h and rhr of the output plane K-2 for D (cache hits), loaded before
        // this step's loads of plane K+1: a wait for them then does not wait
        // for those too (loads complete in issue order for s_waitcnt).  h of
        // plane K-1 for C is the queue's.
        const T hQ = ldo(pl(h, SXY, kc(K - 2)), so);
        T rQ = T(0);
        if constexpr (!DENS) rQ = ldo(pl(rhr, SXY, kc(K - 2)), so);
        const T hP = hq1;
        // Temporaries of a step that are read only under the condition
        // that sets them are left unset rather than zeroed: a default value
        // would cost a register move per step on every path.
        // x of plane K+1 for the LDS: ring cell, and the own cell where it
        // is not in the queue (positions outside the domain)
        T xr, xown;
        if (An_on) {
            if (t < NRING) xr = ld(ring, K + 1);
            if (!in_dom) xown = ld(own, K + 1);
        }
        // neighbour-column inputs of V1/V2 on plane P=K-1.  Column
        // (gi-1, gj) is the source column of the lane to the left wherever
        // V1 is formed (gi >= 2 in the domain, so the left position is in
        // the domain too), and its queues hold h, u2 and u3 of plane P:
        // taken across lanes.  Column (gi, gj-1) is likewise the source
        // column of the position below wherever V2 is formed (gj >= 2): with
        // XJ its h, u1 and u3 go through LDS (written in stage A, read in
        // B); otherwise they are loaded, before stage A so that the latency
        // overlaps it, by every lane on every plane (the plane clamped into
        // the domain; the source column is in the domain, so its j-1
        // neighbour is in the arrays), which costs nothing in a wave that
        // loads anyway.
        const int P = K - 1;
        const bool BC_on = P >= max(1, k0 - 1) && P <= min(lp, k1 + 1);
        const int Pc = kc(P);
        const T nv1_h = from_left_lane(hq1);
        const T nv1_u2a = from_left_lane(u2a0);
        const T nv1_u2b = from_left_lane(u2b0);
        const T nv1_u3a = from_left_lane(u3q0);
        const T nv1_u3b = from_left_lane(u3q1);
        // column (gi, gj-1)
        T nv2_h, nv2_u1a, nv2_u1b, nv2_u3a, nv2_u3b;  // read under BC_on && do_v2
        if constexpr (XJ<RX, RY>::value) {
            if (BC_on) {
                LHI(e_h, 0) = hq1;  // element e = t
                LHI(e_u1, 0) = u1a0;
                LHI(e_u3a, 0) = u3q0;
                LHI(e_u3b, 0) = u3q1;
            }
        } else {
            nv2_h = ldo(pl(h, SXY, Pc) - o.sx, so);
            nv2_u1a = ldo(pl(u1, FXY, Pc) - o.fx, fo);
            nv2_u1b = ldo(pl(u1, FXY, Pc) - o.fx + 1, fo);
            nv2_u3a = ldo(pl(u3, SXY, Pc) - o.sx, so);
            nv2_u3b = ldo(pl(u3, SXY, Pc + 1) - o.sx, so);
        }

        // ---- A: xant on plane K ------------------------------------------
        T xa_new;  // read under A_on, or in C where P < lp implies it
        // A and C are instantiated twice: for the ground and sky planes, and
        // for the interior planes, where the boundary cases compile away
        // (a uniform branch picks one; otherwise both sides of each
        // boundary case would be computed and selected on every plane)
        auto stage_a = [&](auto edge) __attribute__((always_inline)) {
            constexpr bool EDGE = decltype(edge)::value;
            auto& sx = L.sx[E0];
            const bool gnd = EDGE && K == 1, sky = EDGE && K == lp;
            T xc = xq2;
            if constexpr (PASS == 2) {
                xa_new = xaq;
            } else {
            T hinv = rcp_<FAST>(hq2);
            T f1p = donor(xc, sx[saj][sai + 1], u1b1);
            T f1 = donor(sx[saj][sai - 1], xc, u1a1);
            T f2p = donor(xc, sx[saj + 1][sai], u2b1);
            T f2 = donor(sx[saj - 1][sai], xc, u2a1);
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
            if constexpr (DENS) xa_new = (xc - div_<FAST>(f1p - f1 + f2p - f2 + f3p - f3, hq2));
            else xa_new = rq2 * (xc - (f1p - f1 + f2p - f2 + f3p - f3) * hinv);
            }
            sxa[R0][pj][pi] = xa_new;
            // bounds of x for the limiter on plane K: the 7-point stencil,
            // where out-of-domain neighbours are edge copies of the centre
            // (cp_last_to_halo_xy) and leave max/min unchanged
            T m = xc, n = xc;
            if (nxm) { T v = sx[pj + 1][pi]; m = max2(m, v); n = min2(n, v); }
            if (nxp) { T v = sx[pj + 1][pi + 2]; m = max2(m, v); n = min2(n, v); }
            if (nym) { T v = sx[pj][pi + 1]; m = max2(m, v); n = min2(n, v); }
            if (nyp) { T v = sx[pj + 2][pi + 1]; m = max2(m, v); n = min2(n, v); }
            if (!gnd) { m = max2(m, xq1); n = min2(n, xq1); }
            if (!sky) { m = max2(m, xq3); n = min2(n, xq3); }
            xmx1 = m;
            xmn1 = n;
        };
        if (A_on) {
            if (K == 1 || K == lp) stage_a(std::true_type{});
            else stage_a(std::false_type{});
        }
        if (An_on) {
            sx[E1][pj + 1][pi + 1] = in_dom ? xq3 : xown;
            if (t < NRING) sx[E1][rsj][rsi] = xr;
        }
        __syncthreads();
// xant, cp, cn of plane K-d at tile position (a, b)
#define XA(d, a, b) sxa[(d) == 0 ? R0 : ((d) == 1 ? R1 : R2)][(b) + 2][(a) + 2]
#define CP(d, a, b) LMID(scp, ((d) & 1) ? E1 : E0, (b) - lj, (a) - li)
#define CN(d, a, b) LMID(scn, ((d) & 1) ? E1 : E0, (b) - lj, (a) - li)
        if constexpr (XJ<RX, RY>::value) {
            if (BC_on && do_v2) {  // pj >= 1 where V2 is formed
                nv2_h = LHI(e_h, -RX);  // element e = t - RX
                nv2_u1a = LHI(e_u1, -RX);
                nv2_u3a = LHI(e_u3a, -RX);
                nv2_u3b = LHI(e_u3b, -RX);
                // u1 on face i+1 of column j-1: the next position's face i,
                // except on the last column, whose right neighbour is a
                // mirrored ring position
                if (gi < np) nv2_u1b = LHI(e_u1, -RX + 1);
                else nv2_u1b = ldo(pl(u1, FXY, Pc) - o.fx + 1, fo);
            }
        }
        const int Q = K - 2;
        const bool Dq = Q >= k0 && Q <= k1;
        // ---- B: pair sums, V1/V2 on plane P=K-1, V3 on face K --------------
        T s1n, s2n, v3n = T(0);  // s1n/s2n read in B's z terms, where A_on holds
        if (A_on) {
            if (do_v1) s1n = xa_new + XA(0, li - 1, lj);
            if (do_v2) s2n = xa_new + XA(0, li, lj - 1);
            if (do_c && K >= 2 && K >= k0 - 1) {
                T hmz = rcp_<FAST>(T(0.5) * (hq1 + hq2));
                T s1 = u1a0 + u1a1 + u1b1 + u1b0;
                T s2 = u2a0 + u2b0 + u2b1 + u2a1;
                T cr = s1 * rat4(XA(1, li - 1, lj), XA(0, li - 1, lj), XA(1, li + 1, lj),
                                 XA(0, li + 1, lj)) +
                       s2 * rat4(XA(1, li, lj - 1), XA(0, li, lj - 1), XA(1, li, lj + 1),
                                 XA(0, li, lj + 1));
                T a = u3q1;
                v3n = vdyf(XA(1, li, lj), xa_new, a, hmz) - T(0.125) * a * hmz * (cr);
            }
        }
        if (BC_on) {
            const bool zterms = !(P == 1 || P == lp);
            T val1 = T(0), val2 = T(0);
            if (do_v1) {
                T hmx = rcp_<FAST>(T(0.5) * (nv1_h + hq1));
                T s2 = nv1_u2a + nv1_u2b + u2b0 + u2a0;
                T cr = s2 * rat4(XA(1, li - 1, lj - 1), XA(1, li, lj - 1), XA(1, li - 1, lj + 1),
                                 XA(1, li, lj + 1));
                if (zterms) {
                    T s3 = nv1_u3a + nv1_u3b + u3q1 + u3q0;
                    cr = cr + s3 * ((s1n - (XA(2, li, lj) + XA(2, li - 1, lj))) * T(0.25));
                }
                T a = u1a0;
                val1 = vdyf(XA(1, li - 1, lj), XA(1, li, lj), a, hmx) - T(0.125) * a * hmx * (cr);
            }
            if (do_v2) {
                T hmy = rcp_<FAST>(T(0.5) * (nv2_h + hq1));
                T s1 = nv2_u1a + u1a0 + u1b0 + nv2_u1b;
                T cr = s1 * rat4(XA(1, li - 1, lj - 1), XA(1, li - 1, lj), XA(1, li + 1, lj - 1),
                                 XA(1, li + 1, lj));
                if (zterms) {
                    T s3 = nv2_u3a + u3q0 + u3q1 + nv2_u3b;
                    cr = cr + s3 * ((s2n - (XA(2, li, lj) + XA(2, li, lj - 1))) * T(0.25));
                }
                T a = u2a0;
                val2 = vdyf(XA(1, li, lj - 1), XA(1, li, lj), a, hmy) - T(0.125) * a * hmy * (cr);
            }
            sv1[E1][pj][pi] = val1;
            LMID(sv2, E1, 0, 0) = val2;
            if constexpr (TAPE) {
                if (P >= k0 && P <= k1) {
                    if (own_fx) sto(pl(tp->v[0], FXY, P), tfo, val1);
                    if (own_fy) sto(pl(tp->v[1], YXY, P), tso, val2);
                }
            }
        }
        // ---- limited fluxes through the left x-face and lower y-face of this
        //      position on plane Q (V and cp/cn of Q are from the previous
        //      step); stage D reads them for both faces a cell shares, so
        //      each face flux is formed once
        if (Dq && do_f) {
            LMID(sf1, 0, 0, 0) = gflux(sv1[E0][pj][pi], CP(2, li, lj), CN(2, li - 1, lj),
                                CP(2, li - 1, lj), CN(2, li, lj));
            LMID(sf2, 0, 0, 0) = gflux(LMID(sv2, E0, 0, 0), CP(2, li, lj), CN(2, li, lj - 1),
                                CP(2, li, lj - 1), CN(2, li, lj));
        }
        if constexpr (TAPE)
            if (Dq && own_s) sto(pl(tp->psi, SXY, Q), tso, XA(2, li, lj));
        // The loaded source-column queues are not read after this point:
        // the oldest slot of each is loaded with the plane step K+1 needs;
        // the latency is covered by C, D and the next step's barrier.
        {
            // the oldest slot of each queue becomes the newest of step K+1
            const int Kn = K + 1;
            xq1 = ldo(pl(x, SXY, kc(Kn + 1)), so);
            u1a0 = ldo(pl(u1, FXY, kc(Kn)), fo);
            u1b0 = ldo(pl(u1, FXY, kc(Kn)) + 1, fo);
            u2a0 = ldo(pl(u2, YXY, kc(Kn)), so);
            u2b0 = ldo(pl(u2, YXY, kc(Kn)) + o.yx, so);
            u3q0 = ldo(pl(u3, SXY, fc(Kn + 1)), so);
            hq1 = ldo(pl(h, SXY, kc(Kn)), so);
            if constexpr (!DENS && PASS != 2) rq2 = ldo(pl(rhr, SXY, kc(Kn)), so);
            if constexpr (PASS == 2) xaq = ldo(pl(pio->xa, SXY, kc(Kn)), so);
        }
        __syncthreads();
        const bool D_on = do_d && Dq;
        // ---- C: cp, cn on plane P ------------------------------------------
        auto stage_c = [&](auto edge) __attribute__((always_inline)) {
            constexpr bool EDGE = decltype(edge)::value;
            const bool gnd = EDGE && P == 1, sky = EDGE && P == lp;
            T rp = T(0), rn = T(0);
            if (do_c) {
                T mxo = xmx0, mno = xmn0;
                const T xaP = XA(1, li, lj);
                T mx = XA(1, li - 1, lj), mn = mx;
#define ACC(v, lo, hi) { T t_ = (v); lo = min2(lo, t_); hi = max2(hi, t_); }
                ACC(xaP, mn, mx) ACC(XA(1, li + 1, lj), mn, mx)
                ACC(XA(1, li, lj - 1), mn, mx) ACC(XA(1, li, lj + 1), mn, mx)
                if (!gnd) ACC(XA(2, li, lj), mn, mx)
                if (!sky) ACC(xa_new, mn, mx)
#undef ACC
                mx = max2(mx, mxo);
                mn = min2(mn, mno);
                T a1p = sv1[E1][pj][pi + 1], a1 = sv1[E1][pj][pi];
                T a2p = LMID(sv2, E1, 1, 0), a2 = LMID(sv2, E1, 0, 0);
                T a3p, a3;
                if (gnd) { a3p = v3n; a3 = -a3p; }
                else if (sky) { a3 = v3q2; a3p = -a3; }
                else { a3p = v3n; a3 = v3q2; }
                T hc = hP;
                rp = div_<FAST>((mx - xaP) * hc,
                                (pn(a1p) + pp(a1) + pn(a2p) + pp(a2) + pn(a3p) + pp(a3) + ep));
                rn = div_<FAST>((xaP - mn) * hc,
                                (pp(a1p) + pn(a1) + pp(a2p) + pn(a2) + pp(a3p) + pn(a3) + ep));
            }
            LMID(scp, E1, 0, 0) = rp;
            LMID(scn, E1, 0, 0) = rn;
        };
        if (BC_on) {
            if (P == 1 || P == lp) stage_c(std::true_type{});
            else stage_c(std::false_type{});
        }
        // the lower z-face flux of the first output plane of the range is
        // the upper face of plane k0-1, formed here on the step before (the
        // two-plane cp/cn ring no longer holds plane Q-1 once plane Q is
        // output); later planes carry it from the previous plane
        if (Q == k0 - 1 && Q >= 1 && do_d)
            f3c = gflux(v3q2, CP(1, li, lj), CN(2, li, lj), CP(2, li, lj), CN(1, li, lj));
        // ---- D: the new x on plane Q=K-2 (neighbour data of plane Q was
        //      written in the previous step; own cp/cn of Q+1 just now) ----
        if (D_on) {
            const bool gnd = Q == 1, sky = Q == lp;
            // x/y faces from stage B of this step; the lower z-face flux is
            // the upper one of plane Q-1 (the same expression).  On the sky
            // plane C did not run this step, so the ring still holds Q-1.
            const T F1 = LMID(sf1, 0, 0, 0), F1p = LMID(sf1, 0, 0, 1);
            const T F2 = LMID(sf2, 0, 0, 0), F2p = LMID(sf2, 0, 1, 0);
            T F3, F3p;
            if (gnd) {
                const T v = v3q2;
                F3 = gflux(-v, CP(2, li, lj), CN(1, li, lj), CP(1, li, lj), CN(2, li, lj));
                F3p = gflux(v, CP(1, li, lj), CN(2, li, lj), CP(2, li, lj), CN(1, li, lj));
            } else {
                F3 = f3c;
                if (sky) F3p = gflux(-v3q1, CP(3, li, lj), CN(2, li, lj), CP(2, li, lj), CN(3, li, lj));
                else F3p = gflux(v3q2, CP(1, li, lj), CN(2, li, lj), CP(2, li, lj), CN(1, li, lj));
            }
            f3c = F3p;
            T* const xoq = pl(xo, SXY, Q);
            if constexpr (DENS) {
                const T xnew = (XA(2, li, lj) - div_<FAST>(F1p - F1 + F2p - F2 + F3p - F3, hQ));
                sto(xoq, so, xnew);
                const T xi = xnew * hQ;
                T* const xincq = pl(dio->xinc, SXY, Q);
                const T rho_n = ldo(xincq, so);
                sto(xincq, so, xi);
                sto(pl(dio->pexr1, SXY, Q), so, xi * ldo(pl(dio->pexe, SXY, Q), so));
                sto(pl(dio->pexr2, SXY, Q), so, xi);
                sto(pl(dio->rhoadv, SXY, Q), so, rho_n);
                sto(pl(dio->rhr, SXY, Q), so, div_<FAST>(rho_n, xi));
            } else {
                const T r = rQ;
                if constexpr (FAST)  // r*(xant/r - s/h) = xant - r*s/h, to rounding
                    sto(xoq, so, XA(2, li, lj) - r * ((F1p - F1 + F2p - F2 + F3p - F3) * rcp_nr(hQ)));
                else
                    sto(xoq, so, r * (div_<FAST>(XA(2, li, lj), r) -
                                      div_<FAST>(F1p - F1 + F2p - F2 + F3p - F3, hQ)));
                if constexpr (PASS == 1) {
                    // the limited velocities through the low faces of the
                    // cell, and through the wall faces n+1 at the last cells
                    T* const q1 = pl(pio->ul1, FXY, Q);
                    T* const q2 = pl(pio->ul2, YXY, Q);
                    sto(q1, fo, F1);
                    sto(q2, so, F2);
                    sto(pl(pio->ul3, SXY, Q), so, F3);
                    if (gi == np) sto(q1 + 1, fo, F1p);
                    if (gj == mp) sto(q2 + o.yx, so, F2p);
                    if (sky) sto(pl(pio->ul3, SXY, Q + 1), so, F3p);
                }
                if constexpr (TAPE) {
                    // V3 on the lower face Q of the cell (face 1 = -face 2 on
                    // the ground), and on face lp+1 = -face lp on the sky
                    // plane; cp and cn of the cell; x with the inflow values
                    // on the wall halo (the upwind's input) and with edge
                    // copies there (the limiter's bounds)
                    T* const tv = pl(tp->v[2], SXY, Q);
                    sto(tv, so, gnd ? -v3q2 : v3q1);
                    if (sky) sto(tv + SXY, so, -v3q1);
                    sto(pl(tp->cp, SXY, Q), so, CP(2, li, lj));
                    sto(pl(tp->cn, SXY, Q), so, CN(2, li, lj));
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
            }
            if (gi == 1 || gi == np || gj == 1 || gj == mp) {
                const T xe = ldo(pl(x, SXY, Q), so);
                if (gi == 1) sto(xoq - 1, so, xe);
                if (gi == np) sto(xoq + 1, so, xe);
                if (gj == 1) sto(xoq - o.sx, so, xe);
                if (gj == mp) sto(xoq + o.sx, so, xe);
            }
            if constexpr (DENS) {
                // total mass fluxes through the low faces of the cell: the
                // upwind's donor fluxes of plane Q (x of level n, inflow
                // values at the walls) plus the limited fluxes; the wall
                // faces as (donor + inflow part) - outflow part.  These reach
                // the outputs, so the compare-and-select operators of the
                // reference are used (signed zeros as the reference's).
                const T* const xq = pl(x, SXY, Q);
                const T xc = ldo(xq, so);
                const T xl = gi == 1 ? bcx[(gj - 1) + mp * (Q - 1)] : ldo(xq - 1, so);
                const T xb = gj == 1 ? bcy[(gi - 1) + np * (Q - 1)] : ldo(xq - o.sx, so);
                const T* const u1q = pl(u1, FXY, Q);
                const T* const u2q = pl(u2, YXY, Q);
                const T f1ns = mpdata::donor(xl, xc, ldo(u1q, fo));
                const T f2ns = mpdata::donor(xb, xc, ldo(u2q, so));
                T f3ns;
                if (gnd) f3ns = -mpdata::donor(xc, ldo(pl(x, SXY, Q + 1), so), ldo(pl(u3, SXY, Q + 1), so));
                else f3ns = mpdata::donor(ldo(pl(x, SXY, Q - 1), so), xc, ldo(pl(u3, SXY, Q), so));
                sto(pl(dio->m1, FXY, Q), fo, f1ns + F1);
                sto(pl(dio->m2, YXY, Q), so, f2ns + F2);
                sto(pl(dio->m3, SXY, Q), so, f3ns + F3);
                if (gi == np) {
                    const T w = mpdata::donor(xc, bcx[(gj - 1) + mp * ((Q - 1) + lp)], ldo(u1q + 1, fo));
                    const T a = sv1[E0][pj][pi + 1];
                    sto(pl(dio->m1, FXY, Q) + 1, fo,
                        w + mpdata::pp(a) * mpdata::min3(T(1), CP(2, li + 1, lj), CN(2, li, lj)) -
                            mpdata::pn(a) * mpdata::min3(T(1), CP(2, li, lj), CN(2, li + 1, lj)));
                }
                if (gj == mp) {
                    const T w = mpdata::donor(xc, bcy[(gi - 1) + np * ((Q - 1) + lp)], ldo(u2q + o.yx, so));
                    const T a = LMID(sv2, E0, 1, 0);
                    sto(pl(dio->m2, YXY, Q) + o.yx, so,
                        w + mpdata::pp(a) * mpdata::min3(T(1), CP(2, li, lj + 1), CN(2, li, lj)) -
                            mpdata::pn(a) * mpdata::min3(T(1), CP(2, li, lj), CN(2, li, lj + 1)));
                }
                if (sky) {
                    const T w = -f3ns;
                    const T a = -v3q1;
                    sto(pl(dio->m3, SXY, Q + 1), so,
                        w + mpdata::pp(a) * mpdata::min3(T(1), CP(3, li, lj), CN(2, li, lj)) -
                            mpdata::pn(a) * mpdata::min3(T(1), CP(2, li, lj), CN(3, li, lj)));
                }
            }
        }
#undef XA
#undef CP
#undef CN
        // ---- the computed queues move to plane K+1 by the slot rotation; V3
        //      of face K goes into the slot of face K-2 ------------------------
        v3q1 = v3n;
    };
#undef LMID
#undef LHI
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

template <typename T, int RX, int RY, int MINB, bool FAST = false, int PASS = 0>
__global__ void __launch_bounds__(RX * RY, MINB)
k_fused_gauge2(Geo g, Off o, int ntx, int nty, int nchunk, int clen,
               const T* __restrict__ x, const T* __restrict__ u1, const T* __restrict__ u2,
               const T* __restrict__ u3, const T* __restrict__ rhr, const T* __restrict__ h,
               const T* __restrict__ bcx, const T* __restrict__ bcy, T* __restrict__ xo,
               PassIO<T> pio) {
    __shared__ FusedLds<T, RX, RY> L;
    const int ntile = ntx * nty;
    // Work-item order: workgroups are dealt to the XCDs round-robin
    // (blockIdx % NXCD), each with its own L2.  The workgroups of one XCD
    // take a contiguous run of tiles, item by item in rounds, so that at any
    // time they march neighbouring tiles through the same k-chunk and share
    // the halo columns in that L2.  (A split into equal contiguous plane
    // ranges per workgroup balances the tail better but loses this
    // alignment and measured slower.)
    const int nxcd = gridDim.x >= NXCD ? NXCD : 1;
    const int per_xcd = gridDim.x / nxcd;                 // workgroups per XCD
    const int xcd = blockIdx.x % nxcd, slot = blockIdx.x / nxcd;
    const int nwg = per_xcd * nxcd;                       // workgroups that take items
    if (blockIdx.x >= nwg) return;
    const int tiles_per_xcd = (ntile + nxcd - 1) / nxcd;
    for (int r = slot; ; r += per_xcd) {
        const int chunk = r / tiles_per_xcd, tix = r - chunk * tiles_per_xcd;
        if (chunk >= nchunk) break;
        const int tile = xcd * tiles_per_xcd + tix;
        if (tile >= ntile) continue;
        const int k0 = 1 + chunk * clen, k1 = min(g.lp, k0 + clen - 1);
        fused_item<T, RX, RY, false, FAST, PASS>(g, o, ntx, tile, k0, k1, x, u1, u2, u3, rhr, h,
                                                 bcx, bcy, xo, L, nullptr, &pio);
    }
}

// The two-pass step that also records the adjoint's tape (fused_item's
// TAPE), with the work-item order of k_fused_gauge2.
template <typename T, int RX, int RY, int MINB>
__global__ void __launch_bounds__(RX * RY, MINB)
k_fused_gauge_rec(Geo g, Off o, int ntx, int nty, int nchunk, int clen, const T* __restrict__ x,
                  const T* __restrict__ u1, const T* __restrict__ u2, const T* __restrict__ u3,
                  const T* __restrict__ rhr, const T* __restrict__ h, const T* __restrict__ bcx,
                  const T* __restrict__ bcy, T* __restrict__ xo, StdTape<T> tp) {
    __shared__ FusedLds<T, RX, RY> L;
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
        fused_item<T, RX, RY, false, false, 0, true>(g, o, ntx, tile, k0, k1, x, u1, u2, u3, rhr,
                                                     h, bcx, bcy, xo, L, nullptr, nullptr, &tp);
    }
}

// Halo cells of x_out that the step does not define keep their x_in values:
// the z-halo planes and the x/y halo cells outside the edge copies (all x/y
// halo cells of a periodic direction, which the step leaves alone).  Cell q
// of the three halo slabs -> its offset, or -1 for an edge copy (written by
// the step itself).
__device__ __forceinline__ long long untouched_halo_cell(const Geo& g, long long q,
                                                         long long nz_slab, long long ny_slab) {
    const int ih = g.ih;
    const long long nx = g.np + 2 * ih, ny = g.mp + 2 * ih;
    int i, j, k;
    bool copy = true;
    if (q < nz_slab) {
        long long r = q % (nx * ny);
        int layer = (int)(q / (nx * ny));
        k = layer < ih ? 1 - ih + layer : g.lp + 1 + (layer - ih);
        i = (int)(r % nx) + 1 - ih;
        j = (int)(r / nx) + 1 - ih;
    } else if (q < nz_slab + ny_slab) {
        long long r = q - nz_slab;
        i = (int)(r % nx) + 1 - ih;
        long long tt = r / nx;
        int row = (int)(tt % (2 * ih));
        k = (int)(tt / (2 * ih)) + 1;
        j = row < ih ? 1 - ih + row : g.mp + 1 + (row - ih);
        copy = g.py || !((j == 0 || j == g.mp + 1) && i >= 1 && i <= g.np);
    } else {
        long long r = q - nz_slab - ny_slab;
        int col = (int)(r % (2 * ih));
        long long tt = r / (2 * ih);
        j = (int)(tt % g.mp) + 1;
        k = (int)(tt / g.mp) + 1;
        i = col < ih ? 1 - ih + col : g.np + 1 + (col - ih);
        copy = g.px || !(i == 0 || i == g.np + 1);
    }
    return copy ? off_raw<SCAL>(g, i, j, k) : -1;
}

template <typename T>
__global__ void k_copy_untouched_halo(Geo g, const T* __restrict__ x, T* __restrict__ xo) {
    const int ih = g.ih;
    const long long nx = g.np + 2 * ih, ny = g.mp + 2 * ih;
    const long long nz_slab = 2ll * ih * nx * ny;      // k outside 1..lp
    const long long ny_slab = 2ll * ih * nx * g.lp;    // k in 1..lp, j outside 1..mp
    const long long nx_slab = 2ll * ih * g.mp * g.lp;  // k, j inside, i outside 1..np
    const long long total = nz_slab + ny_slab + nx_slab;
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < total;
         q += (long long)gridDim.x * blockDim.x) {
        const long long c = untouched_halo_cell(g, q, nz_slab, ny_slab);
        if (c >= 0) xo[c] = x[c];
    }
}

// Tile configuration (0 = 32x32 lanes at 4 waves per SIMD, the fastest at
// 512x512x256; 1-4 the alternatives in the switch below) and k-chunk
// length; MPDATA_FUSED_CFG / _CHUNK override them for tuning runs (-1: not
// set).
int fused_config() {
    const char* e = getenv("MPDATA_FUSED_CFG");
    return e ? atoi(e) : -1;
}
int fused_chunk_env() {
    const char* e = getenv("MPDATA_FUSED_CHUNK");
    const int c = e ? atoi(e) : -1;
    return c >= 1 ? c : -1;
}

// Workgroups of a persistent launch.  The kernels deal workgroups to the
// NXCD XCDs round-robin and give XCD x the items of its contiguous share
// (items_per_xcd of them), so the grid is a multiple of NXCD with no more
// workgroups per XCD than that share has items (a smaller share would
// leave slots idle while items of the XCD wait for a second round); fewer
// than NXCD items run without the split.
int persistent_grid(long long nitem, long long items_per_xcd, int ncu, int per_cu) {
    if (nitem < NXCD) return (int)std::max(1LL, nitem);
    const long long slots = (long long)ncu * std::max(per_cu, 1);
    const long long per_xcd = std::max(1LL, std::min(items_per_xcd, slots / NXCD));
    return (int)(per_xcd * NXCD);
}

template <int RX, int RY>
long long items_of(const Geo& g, int clen) {
    const long long ntx = (g.np + RX - 5) / (RX - 4), nty = (g.mp + RY - 5) / (RY - 4);
    return ntx * nty * ((g.lp + clen - 1) / clen);
}

// Wall time of a step, in k-march steps of one workgroup: the XCDs work in
// parallel, each on its contiguous share of the tiles, in rounds of per_xcd
// work items, and an item of clen output planes marches clen + kRamp planes
// (two planes of xant and one of V/cp/cn recomputed below the chunk, two
// above it, and the item's set-up).
constexpr int kRamp = 5;
[[maybe_unused]] long long step_cost(long long ntile, int lp, int clen, long long per_xcd) {
    const long long nchunk = (lp + clen - 1) / clen;
    const long long tiles_per_xcd = (ntile + NXCD - 1) / NXCD;
    const long long rounds = (tiles_per_xcd * nchunk + per_xcd - 1) / per_xcd;
    return rounds * (clen + kRamp);
}

template <typename T, int RX, int RY, int MINB>
int per_cu_of() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused_gauge2<T, RX, RY, MINB>, RX * RY, 0) !=
        hipSuccess)
        return 1;
    return n > 0 ? n : 1;
}

// Tiling and k-chunk of a single-domain step.  32x32 tiles recompute the
// fewest ring cells but give one workgroup per CU: they are used when the
// domain gives every CU at least two 64-plane work items, with the chunk
// length (16 planes or more) of least wall time by step_cost (at
// 512x512x256: 128 planes, three rounds of items per XCD against 5.75, so
// six, with 64).  A smaller domain takes 32x16 tiles (two workgroups per
// CU) and the longest chunk of 64 or 32 planes that gives every workgroup
// slot two items, else 16 (measured: on smaller domains more, shorter
// items beat fewer, longer ones; a rounds model over 64..8 planes chose
// 64 at 256x256x128 and was 7 % slower there), and 8 or 4 planes when the
// domain is so small that its items still fit one per CU.
// k-chunk length of least modelled wall time for RX x RY tiles of `kernel`,
// rounds x (clen + 7): the workgroup slots of the device take the work items
// in rounds, and an item marches its clen output planes plus the planes
// recomputed below and above it and its set-up.
template <int RX, int RY, typename K>
int chunk_by_rounds(const Geo& g, K kernel) {
    int dev = 0, ncu = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        ncu < 1)
        ncu = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, RX * RY, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    (void)hipGetLastError();
    const long long slots = (long long)ncu * per_cu;
    long long best = -1;
    int clen = 64;
    for (int c : {64, 32, 16, 8, 4}) {
        const long long rounds = (items_of<RX, RY>(g, c) + slots - 1) / slots;
        const long long cost = rounds * (std::min(c, g.lp) + 7);
        if (best < 0 || cost < best) {
            best = cost;
            clen = c;
        }
    }
    return std::min(clen, g.lp);
}

template <typename T>
void choose_tiling(const Geo& g, int& cfg, int& clen) {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        ncu < 1)
        ncu = 256;
    const int env_cfg = fused_config(), env_clen = fused_chunk_env();
    if (env_cfg >= 0) {
        cfg = env_cfg;
        clen = env_clen > 0 ? env_clen : 64;
    } else if (items_of<32, 32>(g, 64) >= 2LL * ncu) {
        cfg = 0;
        clen = env_clen;
        if (env_clen <= 0) {
            const long long ntile = items_of<32, 32>(g, g.lp);
            const long long per_xcd = std::max(1LL, (long long)ncu * per_cu_of<T, 32, 32, 4>() / NXCD);
            long long best = -1;
            for (int n = 1; n <= g.lp; n++) {
                const int cl = (g.lp + n - 1) / n;
                if (cl < 16) break;
                const long long cost = step_cost(ntile, g.lp, cl, per_xcd);
                if (best < 0 || cost < best) {
                    best = cost;
                    clen = cl;
                }
            }
            if (best < 0) clen = g.lp;
        }
    } else {
        cfg = 1;
        clen = env_clen > 0 ? env_clen : 16;
        if (env_clen <= 0)
            for (int c : {64, 32}) {
                if (items_of<32, 16>(g, c) >= 4LL * ncu) { clen = c; break; }
            }
        // A small domain whose items all get a CU of their own marches
        // clen + kRamp planes in one round: shorter chunks finish sooner as
        // long as every XCD's share still fits one workgroup per CU
        // (measured: 32^3 68 -> 44 us, 59^3 79 -> 45 us per step).
        if (env_clen <= 0 && clen == 16) {
            const long long tiles_per_xcd = (items_of<32, 16>(g, g.lp) + NXCD - 1) / NXCD;
            for (int c : {8, 4}) {
                if (tiles_per_xcd * ((g.lp + c - 1) / c) > std::max(1, ncu / NXCD)) break;
                clen = c;
            }
        }
    }
    if (clen > g.lp) clen = g.lp;
}

// Persistent launch of the fixed-position kernel: as many workgroups as fit
// on the device at once, each looping over (tile, k-chunk) work items.
template <typename T, int RX, int RY, int MINB, bool FAST = false, int PASS = 0>
hipError_t launch2(const Geo& g, const Off& o, int clen, const T* x, T* xout, const T* u1,
                   const T* u2, const T* u3, const T* rhr, const T* h, const T* bcx, const T* bcy,
                   hipStream_t s, const PassIO<T>& pio = PassIO<T>{}) {
    const int ntx = (g.np + RX - 5) / (RX - 4), nty = (g.mp + RY - 5) / (RY - 4);
    const int nchunk = (g.lp + clen - 1) / clen;
    const int nitem = ntx * nty * nchunk;
    int dev = 0, ncu = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, k_fused_gauge2<T, RX, RY, MINB, FAST, PASS>, RX * RY, 0);
    if (e != hipSuccess) return e;
    const int grid = persistent_grid(nitem, (long long)(ntx * nty + NXCD - 1) / NXCD * nchunk, ncu,
                                     per_cu);
    k_fused_gauge2<T, RX, RY, MINB, FAST, PASS><<<grid, RX * RY, 0, s>>>(
        g, o, ntx, nty, nchunk, clen, x, u1, u2, u3, rhr, h, bcx, bcy, xout, pio);
    return hipGetLastError();
}

// Three iterations (upwind + two corrective passes) as two launches: the
// two-pass kernel writes its result to sc.x2 and its limited velocities to
// sc.ul, then the pass kernel advances x2 with them into xout, with the
// limiter bounds from x (time level n).  Tilings 32x32 and 32x16 lanes.
template <typename T, bool FAST>
hipError_t launch_three(const Geo& g, const Off& o, int cfg, int clen, const T* x, T* xout,
                        const T* u1, const T* u2, const T* u3, const T* rhr, const T* h,
                        const T* bcx, const T* bcy, const FusedScratch<T>& sc, hipStream_t s) {
    PassIO<T> p1{sc.ul[0], sc.ul[1], sc.ul[2], nullptr};
    PassIO<T> p2{nullptr, nullptr, nullptr, sc.x2};
    hipError_t e;
    if (cfg == 0) {
        e = launch2<T, 32, 32, 4, FAST, 1>(g, o, clen, x, sc.x2, u1, u2, u3, rhr, h, bcx, bcy, s, p1);
        if (e == hipSuccess)
            e = launch2<T, 32, 32, 4, FAST, 2>(g, o, clen, x, xout, sc.ul[0], sc.ul[1], sc.ul[2],
                                               rhr, h, bcx, bcy, s, p2);
    } else {
        e = launch2<T, 32, 16, 4, FAST, 1>(g, o, clen, x, sc.x2, u1, u2, u3, rhr, h, bcx, bcy, s, p1);
        if (e == hipSuccess)
            e = launch2<T, 32, 16, 4, FAST, 2>(g, o, clen, x, xout, sc.ul[0], sc.ul[1], sc.ul[2],
                                               rhr, h, bcx, bcy, s, p2);
    }
    return e;
}

// The density mode of the fixed-position kernel (DensIO), with the work-item
// order of k_fused_gauge2.
template <typename T, int RX, int RY, int MINB, bool FAST = false>
__global__ void __launch_bounds__(RX * RY, MINB)
k_fused_dens(Geo g, Off o, int ntx, int nty, int nchunk, int clen, const T* __restrict__ x,
             const T* __restrict__ u1, const T* __restrict__ u2, const T* __restrict__ u3,
             const T* __restrict__ h, const T* __restrict__ bcx, const T* __restrict__ bcy,
             T* __restrict__ xo, DensIO<T> dio) {
    __shared__ FusedLds<T, RX, RY> L;
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
        fused_item<T, RX, RY, true, FAST>(g, o, ntx, tile, k0, k1, x, u1, u2, u3, nullptr, h, bcx,
                                          bcy, xo, L, &dio);
    }
}

template <typename T, int RX, int RY, int MINB, bool FAST = false>
hipError_t launch_dens(const Geo& g, const Off& o, int clen, const T* x, T* xout, const T* u1,
                       const T* u2, const T* u3, const T* h, const T* bcx, const T* bcy,
                       const DensIO<T>& d, hipStream_t s) {
    const int ntx = (g.np + RX - 5) / (RX - 4), nty = (g.mp + RY - 5) / (RY - 4);
    const int nchunk = (g.lp + clen - 1) / clen;
    const int nitem = ntx * nty * nchunk;
    int dev = 0, ncu = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu,
                                                         k_fused_dens<T, RX, RY, MINB, FAST>,
                                                         RX * RY, 0);
    if (e != hipSuccess) return e;
    const int grid = persistent_grid(nitem, (long long)(ntx * nty + NXCD - 1) / NXCD * nchunk, ncu,
                                     per_cu);
    k_fused_dens<T, RX, RY, MINB, FAST><<<grid, RX * RY, 0, s>>>(g, o, ntx, nty, nchunk, clen, x,
                                                                 u1, u2, u3, h, bcx, bcy, xout, d);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Batched step: many domains (ensemble members, tracers) of different shapes
// in one persistent launch.  The work items of all members form one list
// (member by member, each chunk-major over its tiles); the items of a
// contiguous run of that list go to the workgroups of one XCD.
// ---------------------------------------------------------------------------
template <typename T>
struct Member {  // device descriptor of one member
    Geo g;
    Off o;
    int ntx, ntile, clen, item0;
    const T *x, *u1, *u2, *u3, *rhr, *h, *bcx, *bcy;
    T* xo;
};

// Scheduling state of a batched step, after the member descriptors in the
// packed table: the run of the work-item list each XCD takes (items
// bnd[x] .. bnd[x+1]-1, split by modelled cost), and per XCD the counter
// from which its workgroups take their next item and the count of its
// workgroups that have finished.  The last workgroup of an XCD to finish
// resets both, so the next launch (or graph replay) starts from zero with
// no memset.
struct BatchSched {
    int bnd[NXCD + 1];
    int next[NXCD];
    int done[NXCD];
};
template <typename T>
size_t sched_offset(int n) {
    return (sizeof(Member<T>) * (size_t)n + 63) & ~size_t(63);
}

// Work items are taken dynamically within each XCD's run (items differ in
// cost: members have different depths, so chunks march different numbers
// of planes), in list order, so the workgroups of an XCD still march
// neighbouring tiles together.  The items left after the run's whole
// rounds (one item per workgroup of the XCD) are each split into up to four
// k-ranges, so that the last round occupies the workgroups; a k-range of a
// chunk gives the same planes as the whole chunk (8 fields of 256x256x128
// on one grid: 2.02 -> 1.85 ms per step; in the statically scheduled
// single-domain kernel the same split measured slower).
template <typename T, int RX, int RY, int MINB, bool FAST = false>
__global__ void __launch_bounds__(RX * RY, MINB)
k_fused_gauge2_batch(const Member<T>* __restrict__ mem, int nmem, int nitem, int shared,
                     BatchSched* __restrict__ sc) {
    __shared__ FusedLds<T, RX, RY> L;
    __shared__ int s_w;
    const int nxcd = gridDim.x >= NXCD ? NXCD : 1;
    const int per_xcd = gridDim.x / nxcd;
    const int xcd = blockIdx.x % nxcd;
    if (blockIdx.x >= per_xcd * nxcd) return;
    int w0, w1;  // this XCD's run
    if (nxcd == 1) {
        w0 = 0;
        w1 = nitem;
    } else if (shared) {
        // members on one grid: equal shares (the items of one (tile, chunk)
        // for all members are consecutive and cost the same)
        const int per = (nitem + NXCD - 1) / NXCD;
        w0 = min(nitem, xcd * per);
        w1 = min(nitem, w0 + per);
    } else {
        w0 = sc->bnd[xcd];
        w1 = sc->bnd[xcd + 1];
    }
    const int nrun = w1 - w0, full = nrun / per_xcd * per_xcd, rem = nrun - full;
    const int split = rem > 0 ? max(1, min(4, per_xcd / rem)) : 1;
    const int ntake = full + rem * split;
    for (;;) {
        if (threadIdx.x == 0) s_w = atomicAdd(&sc->next[xcd], 1);
        __syncthreads();
        const int i = s_w;
        __syncthreads();  // every thread has read s_w before thread 0 writes it again
        if (i >= ntake) break;
        int w = w0 + i, part = 0, nparts = 1;
        if (i >= full) {
            w = w0 + full + (i - full) / split;
            part = (i - full) % split;
            nparts = split;
        }
        int q, l;
        if (shared) {
            // members on one grid (fields sharing the velocity): the items
            // of one (tile, chunk) for all members are consecutive, so one
            // XCD marches them together and reads u, h, rhr once into its L2
            q = w % nmem;
            l = w / nmem;
        } else {
            int lo = 0, hi = nmem - 1;  // last member with item0 <= w
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (mem[mid].item0 <= w) lo = mid;
                else hi = mid - 1;
            }
            q = lo;
            l = w - mem[q].item0;
        }
        const Member<T>& m = mem[q];
        const int chunk = l / m.ntile, tile = l - chunk * m.ntile;
        const int c0 = 1 + chunk * m.clen, c1 = min(m.g.lp, c0 + m.clen - 1);
        const int len = c1 - c0 + 1;
        const int k0 = c0 + len * part / nparts, k1 = c0 + len * (part + 1) / nparts - 1;
        if (k0 > k1) continue;  // a chunk shorter than the split (uniform: every thread)
        fused_item<T, RX, RY, false, FAST>(m.g, m.o, m.ntx, tile, k0, k1, m.x, m.u1, m.u2, m.u3,
                                           m.rhr, m.h, m.bcx, m.bcy, m.xo, L);
    }
    if (threadIdx.x == 0 && atomicAdd(&sc->done[xcd], 1) == per_xcd - 1) {
        atomicExch(&sc->next[xcd], 0);
        atomicExch(&sc->done[xcd], 0);
    }
}

// Halo copies of every member (blockIdx.y = member).
template <typename T>
__global__ void k_copy_untouched_halo_batch(const Member<T>* __restrict__ mem) {
    const Member<T>& m = mem[blockIdx.y];
    const Geo g = m.g;
    const int ih = g.ih;
    const long long nx = g.np + 2 * ih, ny = g.mp + 2 * ih;
    const long long nz_slab = 2ll * ih * nx * ny;
    const long long ny_slab = 2ll * ih * nx * g.lp;
    const long long nx_slab = 2ll * ih * g.mp * g.lp;
    const long long total = nz_slab + ny_slab + nx_slab;
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < total;
         q += (long long)gridDim.x * blockDim.x) {
        const long long c = untouched_halo_cell(g, q, nz_slab, ny_slab);
        if (c >= 0) m.xo[c] = m.x[c];
    }
}

Off make_off(const Geo& g) {
    Off o;
    o.ih = g.ih;
    o.sx = g.np + 2 * g.ih;
    o.sxy = o.sx * (g.mp + 2 * g.ih);
    o.fx = g.np + 1 + 2 * g.ih;
    o.fxy = o.fx * (g.mp + 2 * g.ih);
    o.yx = g.np + 2 * g.ih;
    o.yxy = o.yx * (g.mp + 1 + 2 * g.ih);
    return o;
}

// The density step on the build FAST selects.  32x32 tiles with 64-plane
// chunks whenever they give at least half the CUs a work item: the step is
// not bandwidth-bound, and these re-read the fewest inputs (measured at
// 256x256x128: 1.60e9 B per call against 2.28e9 B with 32x16 tiles and
// 16-plane chunks, in the same 1.12 ms); the single-domain rule otherwise
// and for the tuning overrides.
template <typename T, bool FAST>
hipError_t dens_step(const Geo& g, const T* x, T* xout, const T* u1, const T* u2, const T* u3,
                     const T* h, const T* bcx, const T* bcy, const DensIO<T>& d, hipStream_t s) {
    const Off o = make_off(g);
    int cfg = 0, cl = 64;
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        ncu < 1)
        ncu = 256;
    if (fused_config() >= 0 || fused_chunk_env() > 0 || 2 * items_of<32, 32>(g, 64) < ncu)
        choose_tiling<T>(g, cfg, cl);
    if (cl > g.lp) cl = g.lp;
    if (cfg == 0)
        return launch_dens<T, 32, 32, 4, FAST>(g, o, cl, x, xout, u1, u2, u3, h, bcx, bcy, d, s);
    return launch_dens<T, 32, 16, 4, FAST>(g, o, cl, x, xout, u1, u2, u3, h, bcx, bcy, d, s);
}

template <int RX, int RY>
void tiles_of(const Geo& g, int& ntx, int& nty) {
    ntx = (g.np + RX - 5) / (RX - 4);
    nty = (g.mp + RY - 5) / (RY - 4);
}

template <typename T, int RX, int RY>
int pack_members(const BatchMember<T>* bm, int n, int clen, Member<T>* out) {
    long long item = 0;
    for (int q = 0; q < n; q++) {
        Member<T>& m = out[q];
        m.g = bm[q].g;
        m.o = make_off(m.g);
        int ntx, nty;
        tiles_of<RX, RY>(m.g, ntx, nty);
        m.ntx = ntx;
        m.ntile = ntx * nty;
        m.clen = clen < m.g.lp ? clen : m.g.lp;
        m.item0 = (int)item;
        m.x = bm[q].x; m.u1 = bm[q].u1; m.u2 = bm[q].u2; m.u3 = bm[q].u3;
        m.rhr = bm[q].rhr; m.h = bm[q].h; m.bcx = bm[q].bcx; m.bcy = bm[q].bcy;
        m.xo = bm[q].xo;
        item += (long long)m.ntile * ((m.g.lp + m.clen - 1) / m.clen);
        if (item > (1ll << 30)) return -1;
    }
    // XCD runs of equal modelled cost: an item marches its chunk's planes
    // plus the two recomputed below and above it (fewer at the lids)
    BatchSched* sc = reinterpret_cast<BatchSched*>(reinterpret_cast<char*>(out) +
                                                   sched_offset<T>(n));
    auto cost = [](const Member<T>& m, int chunk) -> long long {
        const int k0 = 1 + chunk * m.clen, k1 = std::min(m.g.lp, k0 + m.clen - 1);
        return (long long)(std::min(m.g.lp, k1 + 2) - std::max(1, k0 - 2) + 1);
    };
    long long total = 0;
    for (int q = 0; q < n; q++) {
        const Member<T>& m = out[q];
        for (int c = 0; c * m.clen < m.g.lp; c++) total += cost(m, c) * m.ntile;
    }
    long long cum = 0;
    int x = 1;
    sc->bnd[0] = 0;
    for (int q = 0; q < n && x < NXCD; q++) {
        const Member<T>& m = out[q];
        for (int c = 0; c * m.clen < m.g.lp && x < NXCD; c++) {
            const long long ci = cost(m, c);
            for (int t = 0; t < m.ntile && x < NXCD; t++) {
                cum += ci;
                while (x < NXCD && cum * NXCD >= total * x)
                    sc->bnd[x++] = m.item0 + c * m.ntile + t + 1;
            }
        }
    }
    for (; x <= NXCD; x++) sc->bnd[x] = (int)item;
    for (int d = 0; d < NXCD; d++) sc->next[d] = sc->done[d] = 0;
    return (int)item;
}

// Tile configurations of the batched kernel: (id, RX, RY, waves per SIMD),
// shared by the exact and the fast build.
#define MPDATA_BATCH_CONFIGS(X) \
    X(0, 32, 16, 2)             \
    X(1, 32, 16, 4)             \
    X(2, 64, 16, 4)             \
    X(3, 16, 32, 2)             \
    X(4, 64, 8, 2)              \
    X(5, 32, 32, 4)

template <typename T, int RX, int RY, int MINB, bool FAST = false>
hipError_t launch_batch(const void* table, int n, int nitem, int shared, hipStream_t s) {
    int dev = 0, ncu = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, k_fused_gauge2_batch<T, RX, RY, MINB, FAST>, RX * RY, 0);
    if (e != hipSuccess) return e;
    const int grid = persistent_grid(nitem, ((long long)nitem + NXCD - 1) / NXCD, ncu, per_cu);
    BatchSched* sc = reinterpret_cast<BatchSched*>(
        reinterpret_cast<char*>(const_cast<void*>(table)) + sched_offset<T>(n));
    k_fused_gauge2_batch<T, RX, RY, MINB, FAST><<<grid, RX * RY, 0, s>>>(
        (const Member<T>*)table, n, nitem, shared, sc);
    return hipGetLastError();
}

}  // namespace

}  // namespace mpdata
