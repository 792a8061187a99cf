// Adjoint of the standard non-oscillatory MPDATA step (2 iterations, open
// walls, FP64): a recorded forward step keeps its intermediates in a tape
// slot (mpdata_adj_record_f64), and mpdata_adj_step_f64 runs the reverse of
// each stage of antidiff3d-standard-gpubc.f90 and upwind3d-gpubc.f90 in
// reverse order:
//   update (:431-459) -> limited donor fluxes (:372-407) -> limiter
//   (:260-347) -> pseudo-velocities and first-order fluxes (:72-202) ->
//   halo fills (mirror of xant, edge copy of x) and upwind (:101-198).
// Each reverse kernel transposes the linearisation the tangent-linear step
// (the stage kernels on dual numbers) computes: every branch is taken as the
// forward value takes it.  Contributions from one face or cell to several
// neighbours are scattered with atomic adds into the adjoint arrays, so the
// adjoint is deterministic up to the order of those additions.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/mpdata_hip.h"
#include "mpdata_ctx.hpp"
#include "mpdata_fused.hpp"
#include "mpdata_stages.hpp"
#include "mpdata_status.hpp"

using namespace mpdata;

namespace {

#define CELL_HEAD                                                  \
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;       \
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;       \
    const int k = 1 + blockIdx.z;                                  \
    if (i > g.np || j > g.mp || k > g.lp) return;                  \
    const bool gnd = k == 1, sky = k == g.lp;                      \
    const int ci_ = i, cj_ = j, ck_ = k;                           \
    const Nb nb = nb_at(g, i, j, k);

// Element offsets around a centre cell (ci_, cj_, ck_) in 32 bits (check_adj
// keeps every array below 2^31 elements): stencil accesses address their
// cell or face as the centre's offset plus small constant multiples of the
// strides, rather than forming a 64-bit offset from the indices each time.
// z-face arrays share the cell arrays' x/y strides.
struct Nb {
    int s, x, y;                    // the centre in cell/z-face, x-face, y-face arrays
    int sx, sxy, fx, fxy, yx, yxy;  // strides
};
__device__ __forceinline__ Nb nb_at(const Geo& g, int i, int j, int k) {
    Nb n;
    n.sx = (int)nx_of<SCAL>(g);
    n.sxy = n.sx * (int)ny_of<SCAL>(g);
    n.fx = (int)nx_of<FACEX>(g);
    n.fxy = n.fx * (int)ny_of<FACEX>(g);
    n.yx = (int)nx_of<FACEY>(g);
    n.yxy = n.yx * (int)ny_of<FACEY>(g);
    const int a = i + g.ih - 1, b = j + g.ih - 1, c = k + g.ih - 1;
    n.s = a + b * n.sx + c * n.sxy;
    n.x = a + b * n.fx + c * n.fxy;
    n.y = a + b * n.yx + c * n.yxy;
    return n;
}
#define S_(a, ii, jj, kk) (a)[nb.s + ((ii) - ci_) + ((jj) - cj_) * nb.sx + ((kk) - ck_) * nb.sxy]
#define X_(a, ii, jj, kk) (a)[nb.x + ((ii) - ci_) + ((jj) - cj_) * nb.fx + ((kk) - ck_) * nb.fxy]
#define Y_(a, ii, jj, kk) (a)[nb.y + ((ii) - ci_) + ((jj) - cj_) * nb.yx + ((kk) - ck_) * nb.yxy]
#define Z_(a, ii, jj, kk) (a)[nb.s + ((ii) - ci_) + ((jj) - cj_) * nb.sx + ((kk) - ck_) * nb.sxy]
// absolute offsets (targets of the LDS flushes, away from any centre)
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
enum { A_PSI = 0, A_X0, A_CP, A_CN, A_PSI2, A_V1, A_V2, A_V3, A_F1, A_F2, A_F3, A_N };
const int A_KIND[A_N] = {SCAL, SCAL, SCAL, SCAL, SCAL, FACEX, FACEY, FACEZ, FACEX, FACEY, FACEZ};

// The arrays that receive scattered contributions (psi, x0, cp, cn, f1..f3)
// are all zero when a reverse step starts: each is cleared by the kernel
// that consumes it (psi alternates between two buffers: the upwind's
// reverse reads psi at its neighbours and clears psi_next, the buffer of the
// next reverse step).  v1..v3 are written before they are read.
struct Adj {
    R *psi, *psi_next, *x0, *cp, *cn, *v1, *v2, *v3, *f1, *f2, *f3;
    const R* xbar;  // adjoint of the recorded step's output
    R uf;           // the velocity adjoint is added to ubar scaled by uf
};

// The two x inputs of a recorded step from x_in in one pass: xb with the
// inflow halo (putbcinxh, mpdataoperators.inc:253-287) for the upwind, x0
// with the edge-copied halo (cp_last_to_halo_xy, :551-584) for the limiter;
// all other cells are x_in.
__global__ void k_prep_x(Geo g, const R* __restrict__ x, const R* __restrict__ bcx,
                         const R* __restrict__ bcy, R* __restrict__ xb, R* __restrict__ x0) {
    const int ih = g.ih;
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1 - ih;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1 - ih;
    const int k = blockIdx.z + 1 - ih;
    if (i > g.np + ih || j > g.mp + ih) return;
    const long long c = off_raw<SCAL>(g, i, j, k);
    const bool kin = k >= 1 && k <= g.lp, iin = i >= 1 && i <= g.np, jin = j >= 1 && j <= g.mp;
    R vb = x[c], v0 = vb;
    if (kin && jin && (i == 0 || i == g.np + 1)) {
        vb = bcx[(j - 1) + (long long)g.mp * ((k - 1) + (long long)g.lp * (i == 0 ? 0 : 1))];
        v0 = x[off_raw<SCAL>(g, i == 0 ? 1 : g.np, j, k)];
    } else if (kin && iin && (j == 0 || j == g.mp + 1)) {
        vb = bcy[(i - 1) + (long long)g.np * ((k - 1) + (long long)g.lp * (j == 0 ? 0 : 1))];
        v0 = x[off_raw<SCAL>(g, i, j == 0 ? 1 : g.mp, k)];
    }
    xb[c] = vb;
    x0[c] = v0;
}

// ---- block-local accumulation ---------------------------------------------
// Adjoint kernels scatter into neighbouring cells and faces.  A block of
// 64x4 cells (one k plane) sums its contributions in LDS regions that cover
// every target of its threads and then adds each region entry to the global
// array with one atomic: far fewer global atomics, and no contention inside
// the block.
constexpr int BX = 64, BY = 4;

// A rolling LDS window of a block over (x, y) planes: plane z sits in slot
// z mod nz, nx x ny entries from (x0, y0).  Blocks march FKB planes k; a
// plane that no later k of the block reaches is added to the global array
// with one atomic per entry and cleared for reuse (flush_plane).
constexpr int FKB = 4;
struct Region {
    int x0, y0, nx, ny, nz, off;
    __device__ int at(int x, int y, int z) const { return off + ((z % nz) * ny + (y - y0)) * nx + (x - x0); }
    __device__ int plane() const { return nx * ny; }
};

template <int KIND>
__device__ void flush_plane(const Geo& g, R* L, const Region& r, int z, R* a, int tid) {
    const int ih = g.ih;
    const int xmax = g.np + ih + (KIND == FACEX), ymax = g.mp + ih + (KIND == FACEY);
    R* const w = L + r.at(r.x0, r.y0, z);
    for (int q = tid; q < r.plane(); q += BX * BY) {
        const R v = w[q];
        if (v == R(0)) continue;
        w[q] = R(0);
        const int x = r.x0 + q % r.nx, y = r.y0 + q / r.nx;
        if (x <= xmax && y <= ymax) atomicAdd(&a[off_raw<KIND>(g, x, y, z)], v);
    }
}

// ---- limited donor fluxes (:372-407), one face ----------------------------
// fb = adjoint of the flux; m/c = upstream/downstream cells.  The limited
// velocity vl = slimv(v, a1, a2, pm, pc) is formed again from the tape's V,
// cp, cn and xant, with the forward's operations (its sign of a zero result
// may differ, which no branch below sees).
__device__ __forceinline__ void adj_face_flux(R* L, R fb, R v, R pm, R pc, R rm, R rc,
                                              R cp_c, R cn_m, R cp_m, R cn_c, int psib_m,
                                              int psib_c, int cpb_c, int cnb_m, int cpb_m,
                                              int cnb_c, R* vbar) {
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
    const R vl = ppv * K1 - pnv * K2;
    const R y1 = pm / rm, y2 = pc / rc;
    // f = donor(y1, y2, vl)
    const R y1b = fb * (vl > R(0) ? vl : R(0));
    const R y2b = fb * (vl < R(0) ? vl : R(0));
    const R vlb = fb * (ind_pos(vl) * y1 + ind_neg(vl) * y2);
    atomicAdd(&L[psib_m], y1b / rm);
    atomicAdd(&L[psib_c], y2b / rc);
    // vl = slimv(v, a1, a2, pm, pc)
    *vbar = vlb * (ind_pos(v) * K1 + ind_neg(v) * K2);
    const R a1b = vlb * (ppv * Pm - pnv * Nc);
    const R a2b = vlb * (ppv * Nm - pnv * Pc);
    // a1 = min3(1, cp_c, cn_m), a2 = min3(1, cp_m, cn_c)
    if (m1 < cn_m) {
        if (!(R(1) < cp_c)) atomicAdd(&L[cpb_c], a1b);
    } else {
        atomicAdd(&L[cnb_m], a1b);
    }
    if (m2 < cn_c) {
        if (!(R(1) < cp_m)) atomicAdd(&L[cpb_m], a2b);
    } else {
        atomicAdd(&L[cnb_c], a2b);
    }
}

// ---- gauge limited fluxes (antidiff3d-gauge-gpubc.f90:405-422), one face --
// F = pp(V)*min3(1, cp_c, cn_m) - pn(V)*min3(1, cp_m, cn_c)
__device__ __forceinline__ void adj_face_gflux(R* L, R fb, R v, R cp_c, R cn_m, R cp_m, R cn_c,
                                               int cpb_c, int cnb_m, int cpb_m, int cnb_c,
                                               R* vbar) {
    const R m1 = R(1) < cp_c ? R(1) : cp_c;
    const R a1 = m1 < cn_m ? m1 : cn_m;
    const R m2 = R(1) < cp_m ? R(1) : cp_m;
    const R a2 = m2 < cn_c ? m2 : cn_c;
    *vbar = fb * (ind_pos(v) * a1 + ind_neg(v) * a2);
    const R a1b = fb * (v > R(0) ? v : R(0));
    const R a2b = -(fb * (v < R(0) ? -v : R(0)));
    if (m1 < cn_m) {
        if (!(R(1) < cp_c)) atomicAdd(&L[cpb_c], a1b);
    } else {
        atomicAdd(&L[cnb_m], a1b);
    }
    if (m2 < cn_c) {
        if (!(R(1) < cp_m)) atomicAdd(&L[cpb_m], a2b);
    } else {
        atomicAdd(&L[cnb_c], a2b);
    }
}

// GAUGE: the gauge variant's fluxes (no xant in the flux), else the
// standard variant's donor fluxes of the limited velocity.
template <bool GAUGE>
__global__ void __launch_bounds__(BX * BY) k_adj_flux(Geo g, const R* __restrict__ v1,
                                                      const R* __restrict__ v2,
                                                      const R* __restrict__ v3,
                                                      const R* __restrict__ cp,
                                                      const R* __restrict__ cn,
                                                      const R* __restrict__ psi,
                                                      const R* __restrict__ rhr,
                                                      const R* __restrict__ h, Adj a) {
    // targets: the cell and its three low neighbours (planes k-1, k)
    constexpr int N = (BX + 1) * (BY + 1) * 2;
    __shared__ R L[3 * N];
    const int tid = threadIdx.y * BX + threadIdx.x;
    const int i0 = 1 + blockIdx.x * BX, j0 = 1 + blockIdx.y * BY;
    const int kb0 = 1 + blockIdx.z * FKB, kb1 = min(g.lp, kb0 + FKB - 1);
    const Region rp{i0 - 1, j0 - 1, BX + 1, BY + 1, 2, 0};
    const Region rcp{i0 - 1, j0 - 1, BX + 1, BY + 1, 2, N};
    const Region rcn{i0 - 1, j0 - 1, BX + 1, BY + 1, 2, 2 * N};
    for (int q = tid; q < 3 * N; q += BX * BY) L[q] = R(0);
    __syncthreads();
    const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
    for (int k = kb0; k <= kb1; k++) {
    if (i <= g.np && j <= g.mp) {
        const int ci_ = i, cj_ = j, ck_ = k;
        const Nb nb = nb_at(g, i, j, k);
        const bool sky = k == g.lp;
        const R pc = S_(psi, i, j, k), rc = S_(rhr, i, j, k);
        // the update's reverse (out = rhr*(psi/rhr - sum*hinv),
        // antidiff3d-standard-gpubc.f90:431-459): psi's adjoint starts at
        // w/rhr and sum's is -(w*hinv), w = rhr*xbar
#define SUMAT(ii, jj, kk) (-((S_(rhr, ii, jj, kk) * S_(a.xbar, ii, jj, kk)) * (R(1) / S_(h, ii, jj, kk))))
        const R wc = rc * S_(a.xbar, i, j, k);
        const R sc = -(wc * (R(1) / S_(h, i, j, k)));
        atomicAdd(&L[rp.at(i, j, k)], wc / rc);
        if (i > 1) {
            const R fb = SUMAT(i - 1, j, k) - sc;
            if constexpr (GAUGE)
                adj_face_gflux(L, fb, X_(v1, i, j, k), S_(cp, i, j, k), S_(cn, i - 1, j, k),
                               S_(cp, i - 1, j, k), S_(cn, i, j, k), rcp.at(i, j, k), rcn.at(i - 1, j, k),
                               rcp.at(i - 1, j, k), rcn.at(i, j, k), &X_(a.v1, i, j, k));
            else
            adj_face_flux(L, fb, X_(v1, i, j, k), S_(psi, i - 1, j, k), pc,
                          S_(rhr, i - 1, j, k), rc, S_(cp, i, j, k), S_(cn, i - 1, j, k),
                          S_(cp, i - 1, j, k), S_(cn, i, j, k), rp.at(i - 1, j, k), rp.at(i, j, k),
                          rcp.at(i, j, k), rcn.at(i - 1, j, k), rcp.at(i - 1, j, k),
                          rcn.at(i, j, k), &X_(a.v1, i, j, k));
        }
        if (j > 1) {
            const R fb = SUMAT(i, j - 1, k) - sc;
            if constexpr (GAUGE)
                adj_face_gflux(L, fb, Y_(v2, i, j, k), S_(cp, i, j, k), S_(cn, i, j - 1, k),
                               S_(cp, i, j - 1, k), S_(cn, i, j, k), rcp.at(i, j, k), rcn.at(i, j - 1, k),
                               rcp.at(i, j - 1, k), rcn.at(i, j, k), &Y_(a.v2, i, j, k));
            else
            adj_face_flux(L, fb, Y_(v2, i, j, k), S_(psi, i, j - 1, k), pc,
                          S_(rhr, i, j - 1, k), rc, S_(cp, i, j, k), S_(cn, i, j - 1, k),
                          S_(cp, i, j - 1, k), S_(cn, i, j, k), rp.at(i, j - 1, k), rp.at(i, j, k),
                          rcp.at(i, j, k), rcn.at(i, j - 1, k), rcp.at(i, j - 1, k),
                          rcn.at(i, j, k), &Y_(a.v2, i, j, k));
        }
        if (k > 1) {
            // the ground cell counts its upper face twice, the sky cell its lower
            const R c1 = k - 1 == 1 ? R(2) : R(1), c2 = sky ? R(2) : R(1);
            const R fb = c1 * SUMAT(i, j, k - 1) - c2 * sc;
            if constexpr (GAUGE)
                adj_face_gflux(L, fb, Z_(v3, i, j, k), S_(cp, i, j, k), S_(cn, i, j, k - 1),
                               S_(cp, i, j, k - 1), S_(cn, i, j, k), rcp.at(i, j, k), rcn.at(i, j, k - 1),
                               rcp.at(i, j, k - 1), rcn.at(i, j, k), &Z_(a.v3, i, j, k));
            else
            adj_face_flux(L, fb, Z_(v3, i, j, k), S_(psi, i, j, k - 1), pc,
                          S_(rhr, i, j, k - 1), rc, S_(cp, i, j, k), S_(cn, i, j, k - 1),
                          S_(cp, i, j, k - 1), S_(cn, i, j, k), rp.at(i, j, k - 1), rp.at(i, j, k),
                          rcp.at(i, j, k), rcn.at(i, j, k - 1), rcp.at(i, j, k - 1),
                          rcn.at(i, j, k), &Z_(a.v3, i, j, k));
        }
    }
    __syncthreads();
    flush_plane<SCAL>(g, L, rp, k - 1, a.psi, tid);
    flush_plane<SCAL>(g, L, rcp, k - 1, a.cp, tid);
    flush_plane<SCAL>(g, L, rcn, k - 1, a.cn, tid);
    __syncthreads();
    }
    flush_plane<SCAL>(g, L, rp, kb1, a.psi, tid);
    flush_plane<SCAL>(g, L, rcp, kb1, a.cp, tid);
    flush_plane<SCAL>(g, L, rcn, kb1, a.cn, tid);
}
#undef SUMAT

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

// GAUGE: the limiter bounds the pseudo-velocities themselves and their
// adjoints go to those of V; else the donor fluxes of V (formed again here)
// and their adjoints go to a.f1..f3.
template <bool GAUGE>
__global__ void __launch_bounds__(BX * BY) k_adj_limiter(Geo g, const R* __restrict__ x0,
                                                         const R* __restrict__ psi,
                                                         const R* __restrict__ h,
                                                         const R* __restrict__ v1,
                                                         const R* __restrict__ v2,
                                                         const R* __restrict__ v3,
                                                         const R* __restrict__ cp,
                                                         const R* __restrict__ cn, Adj a) {
    // targets: psi and x0 within one cell in every direction (planes
    // k-1..k+1), the six faces (x/y faces of plane k, z-faces k and k+1)
    constexpr int NC = (BX + 2) * (BY + 2) * 3;
    constexpr int N1 = (BX + 1) * BY, N2 = BX * (BY + 1), N3 = BX * BY * 2;
    __shared__ R L[2 * NC + N1 + N2 + N3];
    const int tid = threadIdx.y * BX + threadIdx.x;
    const int i0 = 1 + blockIdx.x * BX, j0 = 1 + blockIdx.y * BY;
    const int kb0 = 1 + blockIdx.z * FKB, kb1 = min(g.lp, kb0 + FKB - 1);
    const Region rp{i0 - 1, j0 - 1, BX + 2, BY + 2, 3, 0};
    const Region rx{i0 - 1, j0 - 1, BX + 2, BY + 2, 3, NC};
    const Region r1{i0, j0, BX + 1, BY, 1, 2 * NC};
    const Region r2{i0, j0, BX, BY + 1, 1, 2 * NC + N1};
    const Region r3{i0, j0, BX, BY, 2, 2 * NC + N1 + N2};
    for (int q = tid; q < 2 * NC + N1 + N2 + N3; q += BX * BY) L[q] = R(0);
    __syncthreads();
    const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
    for (int k = kb0; k <= kb1; k++) {
    if (i <= g.np && j <= g.mp) {
        const int ci_ = i, cj_ = j, ck_ = k;
        const Nb nb = nb_at(g, i, j, k);
        const bool gnd = k == 1, sky = k == g.lp;
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
        // the limiter's inputs: the pseudo-velocities (gauge), or their donor
        // fluxes (standard, :182-202), formed again from the tape's V and
        // xant (zero V on the x/y walls gives zero fluxes there)
        R a1p, a1, a2p, a2, a3p, a3;
        if constexpr (GAUGE) {
            a1p = X_(v1, i + 1, j, k);
            a1 = X_(v1, i, j, k);
            a2p = Y_(v2, i, j + 1, k);
            a2 = Y_(v2, i, j, k);
            if (gnd) { a3p = Z_(v3, i, j, k + 1); a3 = -a3p; }
            else if (sky) { a3 = Z_(v3, i, j, k); a3p = -a3; }
            else { a3p = Z_(v3, i, j, k + 1); a3 = Z_(v3, i, j, k); }
        } else {
            const R pc_ = S_(psi, i, j, k);
            a1p = donor(pc_, S_(psi, i + 1, j, k), X_(v1, i + 1, j, k));
            a1 = donor(S_(psi, i - 1, j, k), pc_, X_(v1, i, j, k));
            a2p = donor(pc_, S_(psi, i, j + 1, k), Y_(v2, i, j + 1, k));
            a2 = donor(S_(psi, i, j - 1, k), pc_, Y_(v2, i, j, k));
            if (gnd) { a3p = donor(pc_, S_(psi, i, j, k + 1), Z_(v3, i, j, k + 1)); a3 = -a3p; }
            else if (sky) { a3 = donor(S_(psi, i, j, k - 1), pc_, Z_(v3, i, j, k)); a3p = -a3; }
            else {
                a3p = donor(pc_, S_(psi, i, j, k + 1), Z_(v3, i, j, k + 1));
                a3 = donor(S_(psi, i, j, k - 1), pc_, Z_(v3, i, j, k));
            }
        }
        const R hc = S_(h, i, j, k);
        const R qp = (a1p < R(0) ? -a1p : R(0)) + (a1 > R(0) ? a1 : R(0)) + (a2p < R(0) ? -a2p : R(0)) +
                     (a2 > R(0) ? a2 : R(0)) + (a3p < R(0) ? -a3p : R(0)) + (a3 > R(0) ? a3 : R(0)) + ep;
        const R qn = (a1p > R(0) ? a1p : R(0)) + (a1 < R(0) ? -a1 : R(0)) + (a2p > R(0) ? a2p : R(0)) +
                     (a2 < R(0) ? -a2 : R(0)) + (a3p > R(0) ? a3p : R(0)) + (a3 < R(0) ? -a3 : R(0)) + ep;
        const R iqp = R(1) / qp, iqn = R(1) / qn;
        const R cpb = S_(a.cp, i, j, k), cnb = S_(a.cn, i, j, k);
        const R tpb = cpb * iqp, tnb = cnb * iqn;
        const R qpb = -(cpb * S_(cp, i, j, k) * iqp);
        const R qnb = -(cnb * S_(cn, i, j, k) * iqn);
        // cp = (mx - xc)*hc/qp, cn = (xc - mn)*hc/qn
        const R mxb = tpb * hc, mnb = -(tnb * hc);
        atomicAdd(&L[rp.at(i, j, k)], -(tpb * hc) + tnb * hc);
        atomicAdd(&L[(mx.src ? rp : rx).at(i + mx.di, j + mx.dj, k + mx.dk)], mxb);
        atomicAdd(&L[(mn.src ? rp : rx).at(i + mn.di, j + mn.dj, k + mn.dk)], mnb);
        // denominators: d pn(y) = -[y<0] dy, d pp(y) = [y>0] dy
        const R b1p = -ind_neg(a1p) * qpb + ind_pos(a1p) * qnb;  // a1p enters qp as pn, qn as pp
        const R b1 = ind_pos(a1) * qpb - ind_neg(a1) * qnb;
        const R b2p = -ind_neg(a2p) * qpb + ind_pos(a2p) * qnb;
        const R b2 = ind_pos(a2) * qpb - ind_neg(a2) * qnb;
        const R b3p = -ind_neg(a3p) * qpb + ind_pos(a3p) * qnb;
        const R b3 = ind_pos(a3) * qpb - ind_neg(a3) * qnb;
        // wall faces of x and y are constants after the boundary condition
        if (i < g.np) atomicAdd(&L[r1.at(i + 1, j, k)], b1p);
        if (i > 1) atomicAdd(&L[r1.at(i, j, k)], b1);
        if (j < g.mp) atomicAdd(&L[r2.at(i, j + 1, k)], b2p);
        if (j > 1) atomicAdd(&L[r2.at(i, j, k)], b2);
        if (gnd) atomicAdd(&L[r3.at(i, j, k + 1)], b3p - b3);
        else if (sky) atomicAdd(&L[r3.at(i, j, k)], b3 - b3p);
        else {
            atomicAdd(&L[r3.at(i, j, k + 1)], b3p);
            atomicAdd(&L[r3.at(i, j, k)], b3);
        }
    }
    __syncthreads();
    flush_plane<SCAL>(g, L, rp, k - 1, a.psi, tid);
    flush_plane<SCAL>(g, L, rx, k - 1, a.x0, tid);
    flush_plane<FACEX>(g, L, r1, k, GAUGE ? a.v1 : a.f1, tid);
    flush_plane<FACEY>(g, L, r2, k, GAUGE ? a.v2 : a.f2, tid);
    flush_plane<FACEZ>(g, L, r3, k, GAUGE ? a.v3 : a.f3, tid);
    __syncthreads();
    }
    flush_plane<SCAL>(g, L, rp, kb1, a.psi, tid);
    flush_plane<SCAL>(g, L, rp, kb1 + 1, a.psi, tid);
    flush_plane<SCAL>(g, L, rx, kb1, a.x0, tid);
    flush_plane<SCAL>(g, L, rx, kb1 + 1, a.x0, tid);
    flush_plane<FACEZ>(g, L, r3, kb1 + 1, GAUGE ? a.v3 : a.f3, tid);
    // cp and cn of the block's cells, read above, are cleared for the next
    // reverse step (last, so that no load waits for these stores)
    if (i <= g.np && j <= g.mp)
        for (int k = kb0; k <= kb1; k++) {
            const long long q = off_raw<SCAL>(g, i, j, k);
            a.cp[q] = R(0);
            a.cn[q] = R(0);
        }
}

// ---- pseudo-velocities and first-order fluxes (:72-202) --------------------
// V = vdyfa(x1, x2, a, r) + sum of vcorra(a, b, y1, y2, r) terms; each vcorra
// term carries its four b faces and the four cells of its y1/y2 sums.  The
// adjoints a block's faces scatter land within one cell (psi) or one face (u)
// of the block's cells, so they are summed in LDS (indices below) first and
// the block adds its region to the global arrays once.
constexpr int PBX = BX, PBY = BY, PKB = 4;
constexpr int LP_N = 3 * (PBY + 2) * (PBX + 2);   // psi cells
constexpr int LU_N = 2 * (PBY + 1) * (PBX + 1);   // faces of one direction
constexpr int L_U1 = LP_N, L_U2 = LP_N + LU_N, L_U3 = LP_N + 2 * LU_N, L_N = LP_N + 3 * LU_N;

struct Corr {
    R b, y1, y2, iy2;   // b = sum of four faces; iy2 = 1/y2
    int bf[4];          // LDS slots of those faces
    int cb[4];          // LDS slots of the cells (plus pair 0,1; minus pair 2,3)
    R sg[4];            // signs of the cell values
};

__device__ __forceinline__ void adj_vcorra(R* L, R vb, R a, const Corr& c, R r, R& ab, R uf) {
    // val = -((((0.125*a)*b)*y1)*r / y2)
    const R t1 = R(0.125) * a, t2 = t1 * c.b, t3 = t2 * c.y1, t4 = t3 * r, q = t4 * c.iy2;
    const R qb = -vb;
    const R t4b = qb * c.iy2;
    const R y2b = -(qb * q * c.iy2);
    const R t3b = t4b * r;
    const R y1b = t3b * t2;
    const R t2b = t3b * c.y1;
    const R bb = uf * (t2b * t1);
    ab += R(0.125) * (t2b * c.b);
#pragma unroll
    for (int q2 = 0; q2 < 4; q2++) atomicAdd(&L[c.bf[q2]], bb);
    // y1 = (|c0| + |c1|) - (|c2| + |c3|), y2 = |c0| + |c1| + |c2| + |c3| + ep
    atomicAdd(&L[c.cb[0]], (y1b + y2b) * c.sg[0]);
    atomicAdd(&L[c.cb[1]], (y1b + y2b) * c.sg[1]);
    atomicAdd(&L[c.cb[2]], (-y1b + y2b) * c.sg[2]);
    atomicAdd(&L[c.cb[3]], (-y1b + y2b) * c.sg[3]);
}

// vb: adjoint of V at this face; fb: adjoint of fns = donor(x1, x2, V)
__device__ __forceinline__ void adj_face_pseudo(R* L, R vb, R fb, R V, R x1, R x2, R a, R r,
                                                int x1s, int x2s, int as, const Corr& c0,
                                                const Corr& c1, bool two, R uf) {
    const R ep = Eps<R>::v;
    // fns = donor(x1, x2, V)
    vb += fb * (ind_pos(V) * x1 + ind_neg(V) * x2);
    R x1bv = fb * (V > R(0) ? V : R(0));
    R x2bv = fb * (V < R(0) ? V : R(0));
    // vdyfa: (|a| - a*a*r) * (|x2| - |x1|) / (|x2| + |x1| + ep)
    const R P = fabs(a) - a * a * r, N = fabs(x2) - fabs(x1), D = fabs(x2) + fabs(x1) + ep;
    const R iD = R(1) / D;
    const R val = P * N * iD;
    const R tb = vb * iD, Db = -(vb * val * iD);
    const R Pb = tb * N, Nb = tb * P;
    x2bv += (Nb + Db) * sgn(x2);
    x1bv += (-Nb + Db) * sgn(x1);
    R ab = Pb * sgn(a) - Pb * r * (a + a);
    adj_vcorra(L, vb, a, c0, r, ab, uf);
    if (two) adj_vcorra(L, vb, a, c1, r, ab, uf);
    atomicAdd(&L[x1s], x1bv);
    atomicAdd(&L[x2s], x2bv);
    atomicAdd(&L[as], uf * ab);
}

// gauge corrective term -((0.125*a)*r)*cr, cr = sum of b * rat4 over the
// corr terms; this corr's share (y1 = its rat4 value, plus pair 0,1 and
// minus pair 2,3 with weights +-0.25)
__device__ __forceinline__ void adj_gcorr(R* L, R vb, R a, const Corr& c, R r, R& ab, R uf) {
    const R crb = -(vb * (R(0.125) * a * r));
    const R bb = uf * (crb * c.y1);
#pragma unroll
    for (int q2 = 0; q2 < 4; q2++) atomicAdd(&L[c.bf[q2]], bb);
    const R cw = R(0.25) * c.b * crb;
    atomicAdd(&L[c.cb[0]], cw);
    atomicAdd(&L[c.cb[1]], cw);
    atomicAdd(&L[c.cb[2]], -cw);
    atomicAdd(&L[c.cb[3]], -cw);
    ab -= vb * (R(0.125) * r) * (c.b * c.y1);
}

// gauge pseudo-velocity (antidiff3d-gauge-gpubc.f90:89-203):
// V = (|a| - a*a*r) * ((x2 - x1)*0.5) - ((0.125*a)*r)*cr; vb its adjoint
__device__ __forceinline__ void adj_face_gauge(R* L, R vb, R x1, R x2, R a, R r, int x1s, int x2s,
                                               int as, const Corr& c0, const Corr& c1, bool two,
                                               R uf) {
    const R P = fabs(a) - a * a * r, D = (x2 - x1) * R(0.5);
    const R xb = R(0.5) * P * vb;
    R ab = vb * D * (sgn(a) - r * (a + a));
    adj_gcorr(L, vb, a, c0, r, ab, uf);
    if (two) adj_gcorr(L, vb, a, c1, r, ab, uf);
    atomicAdd(&L[x1s], -xb);
    atomicAdd(&L[x2s], xb);
    atomicAdd(&L[as], uf * ab);
}

// GAUGE: the gauge variant's pseudo-velocities, whose adjoint a.v already
// holds the limiter's part; else the standard variant's, with the adjoint
// a.f of their donor fluxes.
template <bool GAUGE>
__global__ void __launch_bounds__(PBX * PBY) k_adj_pseudo(Geo g, const R* __restrict__ u1,
                                                          const R* __restrict__ u2,
                                                          const R* __restrict__ u3,
                                                          const R* __restrict__ psi,
                                                          const R* __restrict__ h,
                                                          const R* __restrict__ v1,
                                                          const R* __restrict__ v2,
                                                          const R* __restrict__ v3,
                                                          R* __restrict__ ub1, R* __restrict__ ub2,
                                                          R* __restrict__ ub3, Adj a) {
    // A block marches PKB planes k.  LDS windows: psi cells (i0-1..i0+PBX,
    // j0-1..j0+PBY) on planes k-1..k+1, in slot (plane mod 3); u1 faces
    // (i0..i0+PBX, j0-1..j0+PBY-1) and u2 faces (i0-1..i0+PBX-1, j0..j0+PBY)
    // on planes k-1..k, u3 faces (i0-1..i0+PBX-1, j0-1..j0+PBY-1) on k..k+1,
    // in slot (plane mod 2).  After plane k, psi plane k-1, u1/u2 plane k-1
    // and u3 face plane k get no further contributions from this block: they
    // are added to the global arrays and their slots cleared for reuse.
    __shared__ R L[L_N];
    // gauge: psi of the same cells (the tape's xant, halo included) on
    // planes k-1..k+1, slot (plane mod 3): every psi value a face reads,
    // loaded once (measured: 75 -> 69 us per call beside the recording; the
    // standard variant, at 153 VGPRs, ran slower with it and reads psi from
    // memory)
    __shared__ R pv[GAUGE ? LP_N : 1];
    const int tid = threadIdx.y * PBX + threadIdx.x;
    const int i0 = 1 + blockIdx.x * PBX, j0 = 1 + blockIdx.y * PBY;
    const int kb0 = 1 + blockIdx.z * PKB, kb1 = min(g.lp, kb0 + PKB - 1);
    auto fill_psi = [&](int kk) {
        if (!GAUGE || kk < 1 - g.ih || kk > g.lp + g.ih) return;
        R* const w = pv + (kk % 3) * ((PBY + 2) * (PBX + 2));
        for (int q = tid; q < (PBY + 2) * (PBX + 2); q += PBX * PBY) {
            const int ii = i0 - 1 + q % (PBX + 2), jj = j0 - 1 + q / (PBX + 2);
            if (ii <= g.np + g.ih && jj <= g.mp + g.ih) w[q] = *SP(psi, ii, jj, kk);
        }
    };
    for (int q = tid; q < L_N; q += PBX * PBY) L[q] = R(0);
    fill_psi(kb0 - 1);
    fill_psi(kb0);
    fill_psi(kb0 + 1);
    __syncthreads();
    const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
    const bool valid = i <= g.np && j <= g.mp;
    const R ep = Eps<R>::v;
    const int ih = g.ih;
    constexpr int PS = (PBY + 2) * (PBX + 2), US = (PBY + 1) * (PBX + 1);  // one plane
    // adds one plane of a window to the global array and clears it
    auto flush_psi = [&](int kk) {
        R* const w = L + (kk % 3) * PS;
        for (int q = tid; q < PS; q += PBX * PBY) {
            const R v = w[q];
            if (v == R(0)) continue;
            w[q] = R(0);
            const int ii = i0 - 1 + q % (PBX + 2), jj = j0 - 1 + q / (PBX + 2);
            if (ii <= g.np + ih && jj <= g.mp + ih) acc(SP(a.psi, ii, jj, kk), v);
        }
    };
    auto flush_u12 = [&](int kk) {
        R* const w1 = L + L_U1 + (kk % 2) * US;
        R* const w2 = L + L_U2 + (kk % 2) * US;
        for (int q = tid; q < US; q += PBX * PBY) {
            const int a0 = q % (PBX + 1), b0 = q / (PBX + 1);
            R v = w1[q];
            if (v != R(0)) {
                w1[q] = R(0);
                const int ii = i0 + a0, jj = j0 - 1 + b0;
                if (ii <= g.np + 1 + ih && jj <= g.mp + ih) acc(XP(ub1, ii, jj, kk), v);
            }
            v = w2[q];
            if (v != R(0)) {
                w2[q] = R(0);
                const int ii = i0 - 1 + a0, jj = j0 + b0;
                if (ii <= g.np + ih && jj <= g.mp + 1 + ih) acc(YP(ub2, ii, jj, kk), v);
            }
        }
    };
    auto flush_u3 = [&](int kk) {
        R* const w3 = L + L_U3 + (kk % 2) * US;
        for (int q = tid; q < US; q += PBX * PBY) {
            const R v = w3[q];
            if (v == R(0)) continue;
            w3[q] = R(0);
            const int ii = i0 - 1 + q % (PBX + 1), jj = j0 - 1 + q / (PBX + 1);
            if (ii <= g.np + ih && jj <= g.mp + ih) acc(ZP(ub3, ii, jj, kk), v);
        }
    };
    for (int k = kb0; k <= kb1; k++) {
        const bool gnd = k == 1, sky = k == g.lp;
        const bool zterms = !(gnd || sky);
        const int ci_ = i, cj_ = j, ck_ = k;
        const Nb nb = nb_at(g, i, j, k);
#define LPS(ii, jj, kk) (((kk) % 3) * PS + ((jj) - j0 + 1) * (PBX + 2) + ((ii) - i0 + 1))
#define LU1(ii, jj, kk) (L_U1 + ((kk) % 2) * US + ((jj) - j0 + 1) * (PBX + 1) + ((ii) - i0))
#define LU2(ii, jj, kk) (L_U2 + ((kk) % 2) * US + ((jj) - j0) * (PBX + 1) + ((ii) - i0 + 1))
#define LU3(ii, jj, kk) (L_U3 + ((kk) % 2) * US + ((jj) - j0 + 1) * (PBX + 1) + ((ii) - i0 + 1))
    #define AXV(ii, jj, kk) (GAUGE ? pv[GAUGE ? LPS(ii, jj, kk) : 0] : S_(psi, ii, jj, kk))
    #define CORR(c, F1, F2, F3, F4, BF1, BF2, BF3, BF4, I0, J0, K0, I1, J1, K1, I2, J2, K2, I3, J3, K3) \
        {                                                                                          \
            c.bf[0] = BF1; c.bf[1] = BF2; c.bf[2] = BF3; c.bf[3] = BF4;                            \
            c.b = F1 + F2 + F3 + F4;                                                               \
            const R v0_ = AXV(I0, J0, K0), v1_ = AXV(I1, J1, K1);                                  \
            const R v2_ = AXV(I2, J2, K2), v3_ = AXV(I3, J3, K3);                                  \
            c.cb[0] = LPS(I0, J0, K0); c.cb[1] = LPS(I1, J1, K1);                                  \
            c.cb[2] = LPS(I2, J2, K2); c.cb[3] = LPS(I3, J3, K3);                                  \
            if constexpr (GAUGE) {                                                                 \
                c.y1 = ((v1_ + v0_) - (v3_ + v2_)) * R(0.25);                                      \
            } else {                                                                               \
                c.sg[0] = sgn(v0_); c.sg[1] = sgn(v1_); c.sg[2] = sgn(v2_); c.sg[3] = sgn(v3_);    \
                c.y1 = (fabs(v0_) + fabs(v1_)) - (fabs(v2_) + fabs(v3_));                          \
                c.y2 = fabs(v0_) + fabs(v1_) + fabs(v2_) + fabs(v3_) + ep;                         \
                c.iy2 = R(1) / c.y2;                                                               \
            }                                                                                      \
        }
        if (valid && i > 1) {
            const R hm = R(1) / (R(0.5) * (S_(h, i - 1, j, k) + S_(h, i, j, k)));
            Corr c0, c1;
            CORR(c0, Y_(u2, i - 1, j, k), Y_(u2, i - 1, j + 1, k), Y_(u2, i, j + 1, k), Y_(u2, i, j, k),
                 LU2(i - 1, j, k), LU2(i - 1, j + 1, k), LU2(i, j + 1, k), LU2(i, j, k),
                 i - 1, j + 1, k, i, j + 1, k, i - 1, j - 1, k, i, j - 1, k)
            if (zterms)
                CORR(c1, Z_(u3, i - 1, j, k), Z_(u3, i - 1, j, k + 1), Z_(u3, i, j, k + 1), Z_(u3, i, j, k),
                     LU3(i - 1, j, k), LU3(i - 1, j, k + 1), LU3(i, j, k + 1), LU3(i, j, k),
                     i - 1, j, k + 1, i, j, k + 1, i - 1, j, k - 1, i, j, k - 1)
            if constexpr (GAUGE)
                adj_face_gauge(L, X_(a.v1, i, j, k), AXV(i - 1, j, k), AXV(i, j, k), X_(u1, i, j, k), hm,
                               LPS(i - 1, j, k), LPS(i, j, k), LU1(i, j, k), c0, c1, zterms, a.uf);
            else {
                adj_face_pseudo(L, X_(a.v1, i, j, k), X_(a.f1, i, j, k), X_(v1, i, j, k),
                                AXV(i - 1, j, k), AXV(i, j, k), X_(u1, i, j, k), hm,
                                LPS(i - 1, j, k), LPS(i, j, k), LU1(i, j, k), c0, c1, zterms, a.uf);
            }
        }
        if (valid && j > 1) {
            const R hm = R(1) / (R(0.5) * (S_(h, i, j - 1, k) + S_(h, i, j, k)));
            Corr c0, c1;
            CORR(c0, X_(u1, i, j - 1, k), X_(u1, i, j, k), X_(u1, i + 1, j, k), X_(u1, i + 1, j - 1, k),
                 LU1(i, j - 1, k), LU1(i, j, k), LU1(i + 1, j, k), LU1(i + 1, j - 1, k),
                 i + 1, j - 1, k, i + 1, j, k, i - 1, j - 1, k, i - 1, j, k)
            if (zterms)
                CORR(c1, Z_(u3, i, j - 1, k), Z_(u3, i, j, k), Z_(u3, i, j, k + 1), Z_(u3, i, j - 1, k + 1),
                     LU3(i, j - 1, k), LU3(i, j, k), LU3(i, j, k + 1), LU3(i, j - 1, k + 1),
                     i, j - 1, k + 1, i, j, k + 1, i, j - 1, k - 1, i, j, k - 1)
            if constexpr (GAUGE)
                adj_face_gauge(L, Y_(a.v2, i, j, k), AXV(i, j - 1, k), AXV(i, j, k), Y_(u2, i, j, k), hm,
                               LPS(i, j - 1, k), LPS(i, j, k), LU2(i, j, k), c0, c1, zterms, a.uf);
            else {
                adj_face_pseudo(L, Y_(a.v2, i, j, k), Y_(a.f2, i, j, k), Y_(v2, i, j, k),
                                AXV(i, j - 1, k), AXV(i, j, k), Y_(u2, i, j, k), hm,
                                LPS(i, j - 1, k), LPS(i, j, k), LU2(i, j, k), c0, c1, zterms, a.uf);
            }
        }
        if (valid && !gnd) {
            const R hm = R(1) / (R(0.5) * (S_(h, i, j, k - 1) + S_(h, i, j, k)));
            Corr c0, c1;
            CORR(c0, X_(u1, i, j, k - 1), X_(u1, i, j, k), X_(u1, i + 1, j, k), X_(u1, i + 1, j, k - 1),
                 LU1(i, j, k - 1), LU1(i, j, k), LU1(i + 1, j, k), LU1(i + 1, j, k - 1),
                 i + 1, j, k - 1, i + 1, j, k, i - 1, j, k - 1, i - 1, j, k)
            CORR(c1, Y_(u2, i, j, k - 1), Y_(u2, i, j + 1, k - 1), Y_(u2, i, j + 1, k), Y_(u2, i, j, k),
                 LU2(i, j, k - 1), LU2(i, j + 1, k - 1), LU2(i, j + 1, k), LU2(i, j, k),
                 i, j + 1, k - 1, i, j + 1, k, i, j - 1, k - 1, i, j - 1, k)
            if constexpr (GAUGE)
                adj_face_gauge(L, Z_(a.v3, i, j, k), AXV(i, j, k - 1), AXV(i, j, k), Z_(u3, i, j, k), hm,
                               LPS(i, j, k - 1), LPS(i, j, k), LU3(i, j, k), c0, c1, true, a.uf);
            else {
                adj_face_pseudo(L, Z_(a.v3, i, j, k), Z_(a.f3, i, j, k), Z_(v3, i, j, k),
                                AXV(i, j, k - 1), AXV(i, j, k), Z_(u3, i, j, k), hm,
                                LPS(i, j, k - 1), LPS(i, j, k), LU3(i, j, k), c0, c1, true, a.uf);
            }
        }
#undef CORR
#undef AXV
#undef LPS
#undef LU1
#undef LU2
#undef LU3
        __syncthreads();
        flush_psi(k - 1);
        flush_u12(k - 1);
        flush_u3(k);
        // plane k+2 into the slot of plane k-1, whose last reader was this k
        if (k < kb1) fill_psi(k + 2);
        __syncthreads();
    }
    // the planes still open after the last one
    flush_psi(kb1);
    flush_psi(kb1 + 1);
    flush_u12(kb1);
    flush_u3(kb1 + 1);
    // the donor-flux adjoints read above (standard variant) are cleared for
    // the next reverse step, last so that no load waits for these stores
    if constexpr (!GAUGE) {
        if (valid)
            for (int k = kb0; k <= kb1; k++) {
                if (i > 1) a.f1[off_raw<FACEX>(g, i, j, k)] = R(0);
                if (j > 1) a.f2[off_raw<FACEY>(g, i, j, k)] = R(0);
                if (k > 1) a.f3[off_raw<FACEZ>(g, i, j, k)] = R(0);
            }
    }
}

// ---- upwind (upwind3d-gpubc.f90:101-198) and the sum of both uses of x -----
// One thread per cell, UBX x UBY cells of one plane per block.
constexpr int UBX = 64, UBY = 4;
// The reverse of the halo fills (mirror of xant, edge copy of x) is folded
// in: psi_in and x0_in gather the halo cells' adjoints.
// xant = rhr*(x - sum*hinv), fluxes donor(x(m), x(c), u); the x halo holds the
// inflow values (no adjoint).  Each cell takes its low faces (and the last
// cell of a direction its high wall face) for the velocity adjoint.
// The adjoint of xant at interior cell (i,j,k) after the reverse halo fills
// (the forward mirrored xant over the full range, y walls first, then x
// walls; the reverse adds each halo cell's adjoint into its source, x walls
// first): its own value plus the x-wall cell it mirrors to, and on the rows
// that the y walls mirror, the y-wall cell with the x-wall cells that were
// moved into it.
__device__ __forceinline__ R psi_in(const Geo& g, const Nb& nb, int ci_, int cj_, int ck_,
                                    const R* __restrict__ p, int i, int j, int k) {
    const int np = g.np, mp = g.mp;
    auto row = [&](int jj) {
        R v = S_(p, i, jj, k);
        if (i == 2) v += S_(p, 0, jj, k);
        if (i == np - 1) v += S_(p, np + 1, jj, k);
        return v;
    };
    R v = row(j);
    if (j == 2) v += row(0);
    if (j == mp - 1) v += row(mp + 1);
    return v;
}

// The adjoint of the edge-copied x at interior cell (i,j,k): its own value
// plus the wall halo cells that copy it.
__device__ __forceinline__ R x0_in(const Geo& g, const Nb& nb, int ci_, int cj_, int ck_,
                                   const R* __restrict__ p, int i, int j, int k) {
    R v = S_(p, i, j, k);
    if (i == 1) v += S_(p, 0, j, k);
    if (i == g.np) v += S_(p, g.np + 1, j, k);
    if (j == 1) v += S_(p, i, 0, k);
    if (j == g.mp) v += S_(p, i, g.mp + 1, k);
    return v;
}

__device__ __forceinline__ R sbar_of(const Geo& g, const Nb& nb, int ci_, int cj_, int ck_,
                                     const Adj& a, const R* rhr, const R* h, int i, int j,
                                     int k) {
    return -(S_(rhr, i, j, k) * psi_in(g, nb, ci_, cj_, ck_, a.psi, i, j, k) *
             (R(1) / S_(h, i, j, k)));
}

__global__ void __launch_bounds__(UBX * UBY) k_adj_upwind(Geo g, const R* __restrict__ u1,
                                                    const R* __restrict__ u2,
                                                    const R* __restrict__ u3,
                                                    const R* __restrict__ xb,
                                                    const R* __restrict__ rhr,
                                                    const R* __restrict__ h, R* __restrict__ xbar,
                                                    R* __restrict__ ub1, R* __restrict__ ub2,
                                                    R* __restrict__ ub3, Adj a) {
    // sbar of the block's cells and a one-cell ring on plane k, formed once
    // each (the same expression as below, so the same bits) and read by the
    // x and y faces of both neighbours; the z faces form sbar of planes k-1
    // and k+1 themselves
    constexpr int SX = UBX + 2, SY = UBY + 2;
    __shared__ R sbl[SX * SY];
    {
        const int i0 = 1 + blockIdx.x * UBX, j0 = 1 + blockIdx.y * UBY, kk = 1 + blockIdx.z;
        for (int q = threadIdx.y * UBX + threadIdx.x; q < SX * SY; q += UBX * UBY) {
            const int ii = i0 - 1 + q % SX, jj = j0 - 1 + q / SX;
            R v = R(0);
            if (ii >= 1 && ii <= g.np && jj >= 1 && jj <= g.mp) {
                const Nb nb = nb_at(g, ii, jj, kk);
                v = sbar_of(g, nb, ii, jj, kk, a, rhr, h, ii, jj, kk);
            }
            sbl[q] = v;
        }
        __syncthreads();
    }
    CELL_HEAD
    const int np = g.np, mp = g.mp, lp = g.lp;
    // sbar of in-plane cell (ii, jj) from the block's LDS tile
    const int sb0 = (threadIdx.y + 1) * SX + threadIdx.x + 1;
#define SBL(di, dj) sbl[sb0 + (di) + (dj) * SX]
    const R w = S_(rhr, i, j, k) * psi_in(g, nb, ci_, cj_, ck_, a.psi, i, j, k);
    const R sc = -(w * (R(1) / S_(h, i, j, k)));
    R xb_c = w + x0_in(g, nb, ci_, cj_, ck_, a.x0, i, j, k);
    const R xc = S_(xb, i, j, k);
    // x faces i (c downstream) and i+1 (c upstream)
    {
        const R fl = (i > 1 ? SBL(-1, 0) : R(0)) - sc;
        const R fh = sc - (i < np ? SBL(1, 0) : R(0));
        const R ul = X_(u1, i, j, k), uh = X_(u1, i + 1, j, k);
        xb_c += fl * (ul < R(0) ? ul : R(0)) + fh * (uh > R(0) ? uh : R(0));
        X_(ub1, i, j, k) += a.uf * (fl * (ind_pos(ul) * S_(xb, i - 1, j, k) + ind_neg(ul) * xc));
        if (i == np)
            X_(ub1, i + 1, j, k) += a.uf * (fh * (ind_pos(uh) * xc + ind_neg(uh) * S_(xb, i + 1, j, k)));
    }
    {
        const R fl = (j > 1 ? SBL(0, -1) : R(0)) - sc;
        const R fh = sc - (j < mp ? SBL(0, 1) : R(0));
        const R ul = Y_(u2, i, j, k), uh = Y_(u2, i, j + 1, k);
        xb_c += fl * (ul < R(0) ? ul : R(0)) + fh * (uh > R(0) ? uh : R(0));
        Y_(ub2, i, j, k) += a.uf * (fl * (ind_pos(ul) * S_(xb, i, j - 1, k) + ind_neg(ul) * xc));
        if (j == mp)
            Y_(ub2, i, j + 1, k) += a.uf * (fh * (ind_pos(uh) * xc + ind_neg(uh) * S_(xb, i, j + 1, k)));
    }
    {
        // z faces 2..lp; the ground cell counts face 2 twice, the sky cell face lp
        if (!gnd) {
            const R c1 = k - 1 == 1 ? R(2) : R(1), c2 = sky ? R(2) : R(1);
            const R fl = c1 * sbar_of(g, nb, ci_, cj_, ck_, a, rhr, h, i, j, k - 1) - c2 * sc;
            const R ul = Z_(u3, i, j, k);
            xb_c += fl * (ul < R(0) ? ul : R(0));
            Z_(ub3, i, j, k) += a.uf * (fl * (ind_pos(ul) * S_(xb, i, j, k - 1) + ind_neg(ul) * xc));
        }
        if (!sky) {
            const R c1 = gnd ? R(2) : R(1), c2 = k + 1 == lp ? R(2) : R(1);
            const R fh = c1 * sc - c2 * sbar_of(g, nb, ci_, cj_, ck_, a, rhr, h, i, j, k + 1);
            const R uh = Z_(u3, i, j, k + 1);
            xb_c += fh * (uh > R(0) ? uh : R(0));
        }
    }
    S_(xbar, i, j, k) = xb_c;
    // clear the cells of a.x0 read above (each is read by this cell only) and
    // this cell's share of the next reverse step's psi buffer: the cell and
    // the wall halo cells it is the edge of (the only halo cells that
    // receive contributions); last, so that no load waits for these stores
    S_(a.x0, i, j, k) = R(0);
    S_(a.psi_next, i, j, k) = R(0);
    if (i == 1) { S_(a.x0, 0, j, k) = R(0); S_(a.psi_next, 0, j, k) = R(0); }
    if (i == np) { S_(a.x0, np + 1, j, k) = R(0); S_(a.psi_next, np + 1, j, k) = R(0); }
    if (j == 1) { S_(a.x0, i, 0, k) = R(0); S_(a.psi_next, i, 0, k) = R(0); }
    if (j == mp) { S_(a.x0, i, mp + 1, k) = R(0); S_(a.psi_next, i, mp + 1, k) = R(0); }
#undef SBL
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

inline dim3 cblock() { return dim3(UBX, UBY, 1); }
inline dim3 cgrid(const Geo& g) { return dim3((g.np + UBX - 1) / UBX, (g.mp + UBY - 1) / UBY, g.lp); }

int check_adj(mpdata_ctx* c) {
    if (!c) return fail(MPDATA_EINVAL, "null context");
    if (c->real_bytes != 8) return fail(MPDATA_EINVAL, "the adjoint step needs an FP64 context");
    if (c->g.px || c->g.py || c->g.pz)
        return fail(MPDATA_EUNSUPPORTED, "the adjoint step covers open boundaries only");
    if (c->g.ih != 1) return fail(MPDATA_EUNSUPPORTED, "the adjoint step needs ih = 1");
    const long long lim = (1LL << 31) - 1;
    if (len_of<FACEX>(c->g) >= lim || len_of<FACEY>(c->g) >= lim || len_of<FACEZ>(c->g) >= lim)
        return fail(MPDATA_EUNSUPPORTED, "the adjoint step needs arrays below 2^31 elements");
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
        if ((rc = ensure(c, &t.v[d], FACEX + d, 8))) return rc;
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

int mpdata_adj_record_variant_f64(mpdata_ctx* c, int slot, int variant, const double* x_in,
                                  double* x_out, const double* u1, const double* u2,
                                  const double* u3, const double* rhr, const double* h,
                                  const double* bcx, const double* bcy, void* stream) {
    int rc = check_adj(c);
    if (rc) return rc;
    if (variant != MPDATA_STANDARD && variant != MPDATA_GAUGE)
        return fail(MPDATA_EINVAL, "the adjoint covers the standard and gauge variants (variant %d)",
                    variant);
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
    TapeSlot& t = c->tape[slot];
    hipStream_t s = (hipStream_t)stream;
    using St = Stages<R>;
    R *xb = (R*)t.xb, *x0 = (R*)t.x0, *psi = (R*)t.psi, *cp = (R*)t.cp, *cn = (R*)t.cn;
    R* v[3];
    for (int d = 0; d < 3; d++) v[d] = (R*)t.v[d];
    t.variant = variant;
    StdTape<R> tp;
    tp.xb = xb;
    tp.x0 = x0;
    tp.psi = psi;
    tp.cp = cp;
    tp.cn = cn;
    for (int d = 0; d < 3; d++) tp.v[d] = v[d];
    if (variant == MPDATA_GAUGE && c->path != MPDATA_PATH_STAGES) {
        // the fused gauge step, writing the tape as it goes
        HIPCHK(Fused<R>::record(g, x_in, x_out, u1, u2, u3, rhr, h, bcx, bcy, tp, s));
        return MPDATA_OK;
    }
    if (variant == MPDATA_GAUGE) {
        // the gauge step's stage sequence (antidiff3d-gauge-gpubc.f90:71-430):
        // the limiter bounds V itself; the limited fluxes go nowhere
        k_prep_x<<<dim3((g.np + 2 * g.ih + 63) / 64, (g.mp + 2 * g.ih + 3) / 4, g.lp + 2 * g.ih),
                   dim3(64, 4, 1), 0, s>>>(g, x_in, bcx, bcy, xb, x0);
        HIPCHK(hipGetLastError());
        HIPCHK(St::upwind(g, u1, u2, u3, xb, psi, rhr, h, s));
        HIPCHK(St::halo_xy(g, psi, St::MIRROR, 1, s));
        HIPCHK(St::pseudo_vel(g, MPDATA_GAUGE, u1, u2, u3, psi, h, v[0], v[1], v[2], nullptr,
                              nullptr, nullptr, s));
        HIPCHK(St::face_bc(g, v[0], v[1], v[2], s));
        HIPCHK(St::limiter(g, x0, psi, h, v[0], v[1], v[2], cp, cn, s));
        HIPCHK(St::halo_xy(g, cp, St::ZERO, 1, s));
        HIPCHK(St::halo_xy(g, cn, St::ZERO, 1, s));
        HIPCHK(St::gauge_update(g, 1, v[0], v[1], v[2], cp, cn, psi, rhr, h, x_out, nullptr,
                                nullptr, nullptr, s));
        return MPDATA_OK;
    }
    if (c->path != MPDATA_PATH_STAGES) {
        // the fused standard step, writing the tape as it goes
        HIPCHK(FusedStd<R>::record(g, x_in, x_out, u1, u2, u3, rhr, h, bcx, bcy, tp, s));
        return MPDATA_OK;
    }
    // the stage sequence of the standard non-oscillatory step; the donor
    // fluxes of V, the limited velocities and the fluxes go to the context's
    // scratch (the adjoint forms the first two again from the tape)
    R *fns[3], *vl[3], *f[3];
    for (int d = 0; d < 3; d++) {
        if ((rc = ensure(c, &S.fns[d], FACEX + d, 8)) || (rc = ensure(c, &S.ul[d], FACEX + d, 8)) ||
            (rc = ensure(c, &S.f[d], FACEX + d, 8)))
            return rc;
        fns[d] = (R*)S.fns[d];
        vl[d] = (R*)S.ul[d];
        f[d] = (R*)S.f[d];
    }
    k_prep_x<<<dim3((g.np + 2 * g.ih + 63) / 64, (g.mp + 2 * g.ih + 3) / 4, g.lp + 2 * g.ih),
               dim3(64, 4, 1), 0, s>>>(g, x_in, bcx, bcy, xb, x0);
    HIPCHK(hipGetLastError());
    HIPCHK(St::upwind(g, u1, u2, u3, xb, psi, rhr, h, s));
    HIPCHK(St::halo_xy(g, psi, St::MIRROR, 1, s));
    HIPCHK(St::pseudo_vel(g, MPDATA_STANDARD, u1, u2, u3, psi, h, v[0], v[1], v[2], fns[0],
                          fns[1], fns[2], s));
    HIPCHK(St::face_bc(g, v[0], v[1], v[2], s));
    HIPCHK(St::face_bc(g, fns[0], fns[1], fns[2], s));
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

int mpdata_adj_record_f64(mpdata_ctx* c, int slot, const double* x_in, double* x_out,
                          const double* u1, const double* u2, const double* u3,
                          const double* rhr, const double* h, const double* bcx,
                          const double* bcy, void* stream) {
    return mpdata_adj_record_variant_f64(c, slot, MPDATA_STANDARD, x_in, x_out, u1, u2, u3, rhr,
                                         h, bcx, bcy, stream);
}

}  // extern "C"

namespace {

// reverse of the antidiffusive iteration (limited fluxes, limiter,
// pseudo-velocities) of the recorded step in `t`, standard or gauge variant
template <bool GAUGE>
void adj_reverse_antidiff(const Geo& g, const TapeSlot& t, const R* u1, const R* u2, const R* u3,
                          const R* rhr, const R* h, R* ubar1, R* ubar2, R* ubar3, const Adj& a,
                          hipStream_t s) {
    const R *v0 = (const R*)t.v[0], *v1 = (const R*)t.v[1], *v2 = (const R*)t.v[2];
    const R *cp = (const R*)t.cp, *cn = (const R*)t.cn, *psi = (const R*)t.psi;
    const dim3 fgrid((g.np + BX - 1) / BX, (g.mp + BY - 1) / BY, (g.lp + FKB - 1) / FKB);
    const dim3 pgrid((g.np + PBX - 1) / PBX, (g.mp + PBY - 1) / PBY, (g.lp + PKB - 1) / PKB);
    k_adj_flux<GAUGE><<<fgrid, dim3(BX, BY, 1), 0, s>>>(g, v0, v1, v2, cp, cn, psi, rhr, h, a);
    k_adj_limiter<GAUGE><<<fgrid, dim3(BX, BY, 1), 0, s>>>(g, (const R*)t.x0, psi, h, v0, v1, v2,
                                                            cp, cn, a);
    k_adj_pseudo<GAUGE><<<pgrid, dim3(PBX, PBY, 1), 0, s>>>(g, u1, u2, u3, psi, h, v0, v1, v2,
                                                             ubar1, ubar2, ubar3, a);
}

}  // namespace

extern "C" {

int mpdata_adj_step_f64(mpdata_ctx* c, int slot, const double* u1, const double* u2,
                        const double* u3, const double* rhr, const double* h, double* xbar,
                        double* ubar1, double* ubar2, double* ubar3, double ufac, void* stream) {
    int rc = check_adj(c);
    if (rc) return rc;
    if (slot < 0 || slot >= (int)c->tape.size() || !c->tape[slot].xb)
        return fail(MPDATA_EINVAL, "tape slot %d is not reserved", slot);
    if (!u1 || !u2 || !u3 || !rhr || !h || !xbar || !ubar1 || !ubar2 || !ubar3)
        return fail(MPDATA_EINVAL, "null pointer");
    const Geo& g = c->g;
    for (int q = 0; q < A_N; q++)
        if ((rc = ensure(c, &c->adj[q], A_KIND[q], 8))) return rc;
    // psi alternates between A_PSI and A_PSI2 from one reverse step to the
    // next (the upwind's reverse clears the other buffer)
    const int pcur = c->adj_psi ? A_PSI2 : A_PSI, pnext = c->adj_psi ? A_PSI : A_PSI2;
    Adj a{(R*)c->adj[pcur], (R*)c->adj[pnext], (R*)c->adj[A_X0], (R*)c->adj[A_CP],
          (R*)c->adj[A_CN], (R*)c->adj[A_V1], (R*)c->adj[A_V2], (R*)c->adj[A_V3],
          (R*)c->adj[A_F1], (R*)c->adj[A_F2], (R*)c->adj[A_F3], xbar, ufac};
    const TapeSlot& t = c->tape[slot];
    hipStream_t s = (hipStream_t)stream;
    HIPCHK(hipGetLastError());
    if (t.variant == MPDATA_GAUGE)
        adj_reverse_antidiff<true>(g, t, u1, u2, u3, rhr, h, ubar1, ubar2, ubar3, a, s);
    else
        adj_reverse_antidiff<false>(g, t, u1, u2, u3, rhr, h, ubar1, ubar2, ubar3, a, s);
    k_adj_upwind<<<cgrid(g), cblock(), 0, s>>>(g, u1, u2, u3, (const R*)t.xb, rhr, h, xbar, ubar1,
                                               ubar2, ubar3, a);
    HIPCHK(hipGetLastError());
    c->adj_psi ^= 1;
    return MPDATA_OK;
}

}  // extern "C"
