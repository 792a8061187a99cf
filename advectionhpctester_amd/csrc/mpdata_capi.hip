// C ABI of libmpdata_hip.so (declared in include/mpdata_hip.h): argument
// checking, per-domain scratch, and the sequencing of the per-stage kernels
// and the fused step kernel on the caller's stream.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mpdata_hip.h"
#include "mpdata_ctx.hpp"
#include "mpdata_dual.hpp"
#include "mpdata_fused.hpp"
#include "mpdata_stages.hpp"
#include "mpdata_status.hpp"
#include "mpdata_timers.hpp"

using namespace mpdata;

namespace mpdata {

namespace {
thread_local std::string g_last_error;
}  // namespace

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int hip_fail(hipError_t e, const char* where) {
    return fail(MPDATA_EHIP, "%s: %s", where, hipGetErrorString(e));
}

const char* last_error() { return g_last_error.c_str(); }

long long field_len(const Geo& g, int kind) {
    switch (kind) {
        case FACEX: return len_of<FACEX>(g);
        case FACEY: return len_of<FACEY>(g);
        case FACEZ: return len_of<FACEZ>(g);
        default: return len_of<SCAL>(g);
    }
}

int ensure(mpdata_ctx* c, void** slot, int kind, int elem_bytes) {
    if (*slot) return MPDATA_OK;
    size_t bytes = (size_t)field_len(c->g, kind) * elem_bytes;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(MPDATA_ENOMEM, "hipMalloc(%zu bytes) for scratch failed: %s", bytes,
                    hipGetErrorString(e));
    }
    // zero once so that halo cells never read are deterministic
    e = hipMemset(p, 0, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMemset(scratch)");
    *slot = p;
    c->owned.push_back(p);
    c->held += (long long)bytes;
    return MPDATA_OK;
}


int check_ctx(const mpdata_ctx* c, int real_bytes) {
    if (!c) return fail(MPDATA_EINVAL, "null context");
    if (c->real_bytes != real_bytes)
        return fail(MPDATA_EINVAL, "context holds %d-byte reals, call needs %d", c->real_bytes,
                    real_bytes);
    int dev = -1;
    if (hipGetDevice(&dev) == hipSuccess && dev != c->device)
        return fail(MPDATA_EINVAL, "context was created on device %d, current device is %d",
                    c->device, dev);
    return MPDATA_OK;
}

int check_dims(const mpdata_ctx* c, int ipoles, int ibcx, int ibcy, int ibcz, int np, int mp,
               int lp, int ih) {
    if (ipoles != 0)
        return fail(MPDATA_EUNSUPPORTED, "ipoles0=%d: only the Cartesian case (0) exists", ipoles);
    const Geo& g = c->g;
    if (np != g.np || mp != g.mp || lp != g.lp || ih != g.ih)
        return fail(MPDATA_EINVAL, "shape (%d,%d,%d,ih=%d) differs from the context (%d,%d,%d,ih=%d)",
                    np, mp, lp, ih, g.np, g.mp, g.lp, g.ih);
    if (ibcx != g.px || ibcy != g.py || ibcz != g.pz)
        return fail(MPDATA_EINVAL, "boundary flags (%d,%d,%d) differ from the context (%d,%d,%d)",
                    ibcx, ibcy, ibcz, g.px, g.py, g.pz);
    return MPDATA_OK;
}

}  // namespace mpdata


namespace {

// Scratch set of a real type: the tangent-linear step keeps its own.
template <typename T>
Scratch& scr(mpdata_ctx* c) {
    return is_dual<T>::value ? c->dual : c->prim;
}

int ensure_red(mpdata_ctx* c, long long bytes) {
    if (c->red_bytes >= bytes) return MPDATA_OK;
    if (c->red) {
        (void)hipFree(c->red);
        c->held -= c->red_bytes;
        c->red = nullptr;
        c->red_bytes = 0;
    }
    hipError_t e = hipMalloc(&c->red, (size_t)bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(MPDATA_ENOMEM, "hipMalloc(%lld bytes) for reductions failed", bytes);
    }
    c->red_bytes = bytes;
    c->held += bytes;
    return MPDATA_OK;
}

// One corrective pass on the per-stage kernels, in the reference's stage
// order (antidiff3d-gauge-gpubc.f90:71-430,
// antidiff3d-standard-gpubc.f90:59-462).
template <typename T>
int stage_pass(mpdata_ctx* c, int variant, int nonos, const T* u1, const T* u2, const T* u3,
               T* psi, T* x0, T* out, const T* rhr, const T* h, T* ul1, T* ul2, T* ul3,
               hipStream_t s) {
    Scratch& S = scr<T>(c);
    using St = Stages<T>;
    const Geo& g = c->g;
    const bool stdv = variant == MPDATA_STANDARD;
    int rc;
    for (int d = 0; d < 3; d++)
        if ((rc = ensure(c, &S.v[d], FACEX + d, sizeof(T)))) return rc;
    if (nonos) {
        if ((rc = ensure(c, &S.cp, SCAL, sizeof(T)))) return rc;
        if ((rc = ensure(c, &S.cn, SCAL, sizeof(T)))) return rc;
    }
    if (stdv)
        for (int d = 0; d < 3; d++) {
            if ((rc = ensure(c, &S.fns[d], FACEX + d, sizeof(T)))) return rc;
            if ((rc = ensure(c, &S.f[d], FACEX + d, sizeof(T)))) return rc;
        }
    T *v1 = (T*)S.v[0], *v2 = (T*)S.v[1], *v3 = (T*)S.v[2];
    T *f1ns = (T*)S.fns[0], *f2ns = (T*)S.fns[1], *f3ns = (T*)S.fns[2];
    T *cp = (T*)S.cp, *cn = (T*)S.cn;
    HIPCHK(St::halo_xy(g, psi, St::MIRROR, 1, s));
    HIPCHK(St::pseudo_vel(g, variant, u1, u2, u3, psi, h, v1, v2, v3, f1ns, f2ns, f3ns, s));
    HIPCHK(St::face_bc(g, v1, v2, v3, s));
    if (stdv) HIPCHK(St::face_bc(g, f1ns, f2ns, f3ns, s));
    HIPCHK(St::halo_xy(g, x0, St::COPY, 0, s));
    if (nonos) {
        HIPCHK(St::limiter(g, x0, psi, h, stdv ? f1ns : v1, stdv ? f2ns : v2, stdv ? f3ns : v3,
                           cp, cn, s));
        HIPCHK(St::halo_xy(g, cp, St::ZERO, 1, s));
        HIPCHK(St::halo_xy(g, cn, St::ZERO, 1, s));
    }
    if (!stdv) {
        HIPCHK(St::gauge_update(g, nonos, v1, v2, v3, cp, cn, psi, rhr, h, out, ul1, ul2, ul3, s));
        return MPDATA_OK;
    }
    T *f1 = (T*)S.f[0], *f2 = (T*)S.f[1], *f3 = (T*)S.f[2];
    HIPCHK(St::std_flux(g, nonos, v1, v2, v3, cp, cn, psi, rhr, f1, f2, f3, ul1, ul2, ul3, s));
    HIPCHK(St::face_bc(g, f1, f2, f3, s));
    if (ul1) HIPCHK(St::face_bc(g, ul1, ul2, ul3, s));
    HIPCHK(St::std_update(g, f1, f2, f3, psi, rhr, h, out, s));
    return MPDATA_OK;
}

template <typename T>
int upwind_impl(mpdata_ctx* c, const T* u1, const T* u2, const T* u3, T* x, T* xant,
                const T* rhr, const T* h, const T* bcx, const T* bcy, hipStream_t s) {
    const Geo& g = c->g;
    if ((!g.px && !bcx) || (!g.py && !bcy))
        return fail(MPDATA_EINVAL, "open boundaries need bcx/bcy");
    HIPCHK(Stages<T>::putbcinxh(g, x, bcx, bcy, s));
    HIPCHK(Stages<T>::upwind(g, u1, u2, u3, x, xant, rhr, h, s));
    return MPDATA_OK;
}

template <typename T>
int step_stages(mpdata_ctx* c, int variant, int iters, int nonos, T* x, const T* u1,
                const T* u2, const T* u3, const T* rhr, const T* h, const T* bcx, const T* bcy,
                hipStream_t s) {
    Scratch& S = scr<T>(c);
    int rc;
    if ((rc = ensure(c, &S.xant, SCAL, sizeof(T)))) return rc;
    T* xant = (T*)S.xant;
    tt_beg(16, s);  // upwind3d-gpubc.f90:87-200
    rc = upwind_impl<T>(c, u1, u2, u3, x, xant, rhr, h, bcx, bcy, s);
    tt_end(16, s);
    if (rc) return rc;
    if (variant == MPDATA_UPWIND || iters <= 1) {
        HIPCHK(Stages<T>::copy_interior(c->g, xant, x, s));
        return MPDATA_OK;
    }
    if (iters > 2) {
        if ((rc = ensure(c, &S.psi2, SCAL, sizeof(T)))) return rc;
        for (int d = 0; d < 3; d++) {
            if ((rc = ensure(c, &S.ul[d], FACEX + d, sizeof(T)))) return rc;
            if ((rc = ensure(c, &S.adv[d], FACEX + d, sizeof(T)))) return rc;
        }
    }
    T* psi = xant;
    T* nxt = (T*)S.psi2;
    T* adv[3] = {(T*)S.adv[0], (T*)S.adv[1], (T*)S.adv[2]};
    T* lim[3] = {(T*)S.ul[0], (T*)S.ul[1], (T*)S.ul[2]};
    for (int it = 2; it <= iters; it++) {
        const bool last = it == iters;
        const T* a1 = it == 2 ? u1 : adv[0];
        const T* a2 = it == 2 ? u2 : adv[1];
        const T* a3 = it == 2 ? u3 : adv[2];
        // antidiff3d-gauge-gpubc.f90:62-434 / antidiff3d-standard-gpubc.f90:54-462
        const int tid = variant == MPDATA_GAUGE ? 51 : 53;
        tt_beg(tid, s);
        rc = stage_pass<T>(c, variant, nonos, a1, a2, a3, psi, x, last ? x : nxt, rhr, h,
                           last ? nullptr : lim[0], last ? nullptr : lim[1],
                           last ? nullptr : lim[2], s);
        tt_end(tid, s);
        if (rc) return rc;
        if (!last) {
            for (int d = 0; d < 3; d++) {
                T* t = adv[d];
                adv[d] = lim[d];
                lim[d] = t;
            }
            T* t = psi;
            psi = nxt;
            nxt = t;
        }
    }
    return MPDATA_OK;
}

int check_step_opts(int variant, int iters, int nonos) {
    if (variant < MPDATA_UPWIND || variant > MPDATA_STANDARD)
        return fail(MPDATA_EINVAL, "variant %d", variant);
    if (variant != MPDATA_UPWIND && (iters < 2 || iters > 8))
        return fail(MPDATA_EINVAL, "iters %d outside 2..8", iters);
    if (nonos != 0 && nonos != 1) return fail(MPDATA_EINVAL, "nonos %d", nonos);
    return MPDATA_OK;
}

// Whether this call runs the fused kernel (after the path checks).
template <typename T>
int use_fused(mpdata_ctx* c, int variant, int iters, int nonos, bool* fused) {
    const bool fusable = Fused<T>::supports(c->g, variant, iters, nonos);
    if ((c->path == MPDATA_PATH_FUSED || c->path == MPDATA_PATH_FAST) && !fusable)
        return fail(MPDATA_EUNSUPPORTED,
                    "fused path supports variant %d iters %d nonos %d only for %s", variant,
                    iters, nonos, Fused<T>::scope());
    *fused = c->path != MPDATA_PATH_STAGES && fusable;
    return MPDATA_OK;
}

// One fused step out of place: under MPDATA_PATH_FAST on the fast build of
// the variant's kernel, else on the exactly rounded one.  Three iterations
// keep the second iterate and the second pass's limited velocities in the
// context's scratch.
template <typename T>
int fused_step(mpdata_ctx* c, int variant, int iters, int nonos, const T* x, T* xout,
               const T* u1, const T* u2, const T* u3, const T* rhr, const T* h, const T* bcx,
               const T* bcy, hipStream_t s) {
    if (variant == MPDATA_STANDARD) {
        const hipError_t e =
            c->path == MPDATA_PATH_FAST
                ? FusedStdFast<T>::step(c->g, nonos, x, xout, u1, u2, u3, rhr, h, bcx, bcy, s)
                : FusedStd<T>::step(c->g, nonos, x, xout, u1, u2, u3, rhr, h, bcx, bcy, s);
        if (e != hipSuccess) return hip_fail(e, "fused standard step");
        return MPDATA_OK;
    }
    FusedScratch<T> sc;
    if (iters == 3) {
        Scratch& S = scr<T>(c);
        int rc;
        if ((rc = ensure(c, &S.psi2, SCAL, sizeof(T))) ||
            (rc = ensure(c, &S.ul[0], FACEX, sizeof(T))) ||
            (rc = ensure(c, &S.ul[1], FACEY, sizeof(T))) ||
            (rc = ensure(c, &S.ul[2], FACEZ, sizeof(T))))
            return rc;
        sc.x2 = (T*)S.psi2;
        for (int d = 0; d < 3; d++) sc.ul[d] = (T*)S.ul[d];
    }
    const hipError_t e =
        c->path == MPDATA_PATH_FAST
            ? FusedFast<T>::step(c->g, x, xout, u1, u2, u3, rhr, h, bcx, bcy, s, iters, &sc)
            : Fused<T>::step(c->g, x, xout, u1, u2, u3, rhr, h, bcx, bcy, s, iters, &sc);
    if (e != hipSuccess) return hip_fail(e, "fused step");
    return MPDATA_OK;
}

template <typename T>
int check_step_args(mpdata_ctx* c, const T* x, const T* u1, const T* u2, const T* u3,
                    const T* rhr, const T* h, const T* bcx, const T* bcy) {
    if (any_null(x, u1, u2, u3, rhr, h)) return fail(MPDATA_EINVAL, "null field pointer");
    if ((!c->g.px && !bcx) || (!c->g.py && !bcy))
        return fail(MPDATA_EINVAL, "open boundaries need bcx/bcy");
    return MPDATA_OK;
}

// In-place step: x <- MPDATA(x).  The fused kernel works out of place into
// the context's buffer, which is then copied back.
template <typename T>
int step_impl(mpdata_ctx* c, int variant, int iters, int nonos, T* x, const T* u1, const T* u2,
              const T* u3, const T* rhr, const T* h, const T* bcx, const T* bcy,
              hipStream_t s) {
    Scratch& S = scr<T>(c);
    int rc;
    bool fused = false;
    if ((rc = check_step_opts(variant, iters, nonos))) return rc;
    if ((rc = check_step_args<T>(c, x, u1, u2, u3, rhr, h, bcx, bcy))) return rc;
    if ((rc = use_fused<T>(c, variant, iters, nonos, &fused))) return rc;
    if (!fused)
        return step_stages<T>(c, variant, iters, nonos, x, u1, u2, u3, rhr, h, bcx, bcy, s);
    if ((rc = ensure(c, &S.xnext, SCAL, sizeof(T)))) return rc;
    T* xn = (T*)S.xnext;
    if ((rc = fused_step<T>(c, variant, iters, nonos, x, xn, u1, u2, u3, rhr, h, bcx, bcy, s))) return rc;
    HIPCHK(hipMemcpyAsync(x, xn, (size_t)len_of<SCAL>(c->g) * sizeof(T), hipMemcpyDeviceToDevice, s));
    return MPDATA_OK;
}

// Out-of-place step: x_out <- MPDATA(x_in); x_in is not modified.
template <typename T>
int step_oop_impl(mpdata_ctx* c, int variant, int iters, int nonos, const T* x_in, T* x_out,
                  const T* u1, const T* u2, const T* u3, const T* rhr, const T* h,
                  const T* bcx, const T* bcy, hipStream_t s) {
    int rc;
    bool fused = false;
    if ((rc = check_step_opts(variant, iters, nonos))) return rc;
    if ((rc = check_step_args<T>(c, x_in, u1, u2, u3, rhr, h, bcx, bcy))) return rc;
    if (!x_out) return fail(MPDATA_EINVAL, "null x_out");
    const size_t bytes = (size_t)len_of<SCAL>(c->g) * sizeof(T);
    const char *a = (const char*)x_in, *b = (const char*)x_out;
    if (a < b + bytes && b < a + bytes) return fail(MPDATA_EINVAL, "x_in and x_out overlap");
    if ((rc = use_fused<T>(c, variant, iters, nonos, &fused))) return rc;
    if (fused) {
        return fused_step<T>(c, variant, iters, nonos, x_in, x_out, u1, u2, u3, rhr, h, bcx, bcy, s);
    }
    HIPCHK(hipMemcpyAsync(x_out, x_in, bytes, hipMemcpyDeviceToDevice, s));
    return step_stages<T>(c, variant, iters, nonos, x_out, u1, u2, u3, rhr, h, bcx, bcy, s);
}

// nt steps.  With the fused kernel the steps alternate between x and the
// context's buffer; an odd count ends with one copy back into x.
template <typename T>
int timeloop_impl(mpdata_ctx* c, int variant, int iters, int nonos, int nt, T* x, const T* u1,
                  const T* u2, const T* u3, const T* rhr, const T* h, const T* bcx,
                  const T* bcy, hipStream_t s) {
    Scratch& S = scr<T>(c);
    int rc;
    bool fused = false;
    if (nt < 0) return fail(MPDATA_EINVAL, "nt %d", nt);
    if ((rc = check_step_opts(variant, iters, nonos))) return rc;
    if ((rc = check_step_args<T>(c, x, u1, u2, u3, rhr, h, bcx, bcy))) return rc;
    if ((rc = use_fused<T>(c, variant, iters, nonos, &fused))) return rc;
    if (!fused) {
        for (int it = 0; it < nt; it++)
            if ((rc = step_stages<T>(c, variant, iters, nonos, x, u1, u2, u3, rhr, h, bcx, bcy, s)))
                return rc;
        return MPDATA_OK;
    }
    if ((rc = ensure(c, &S.xnext, SCAL, sizeof(T)))) return rc;
    T* buf[2] = {x, (T*)S.xnext};
    for (int it = 0; it < nt; it++)
        if ((rc = fused_step<T>(c, variant, iters, nonos, buf[it & 1], buf[(it + 1) & 1], u1, u2,
                                u3, rhr, h, bcx, bcy, s)))
            return rc;
    if (nt & 1)
        HIPCHK(hipMemcpyAsync(x, buf[1], (size_t)len_of<SCAL>(c->g) * sizeof(T),
                              hipMemcpyDeviceToDevice, s));
    return MPDATA_OK;
}

}  // namespace

namespace mpdata {

template <typename T>
int reduce_courants(mpdata_ctx* c, const T* ox, const T* oy, const T* oz, const T* h, int level,
                    T dxi, T dyi, T dzi, T dt, double* out, hipStream_t s) {
    if (any_null(ox, oy, oz, out) || (level == 2 && !h))
        return fail(MPDATA_EINVAL, "null pointer");
    if (level != 1 && level != 2)
        return fail(MPDATA_EINVAL, "ntimelevel %d (reference accepts 1 or 2)", level);
    const long long nb = Stages<T>::reduce_blocks(c->g);
    int rc;
    if ((rc = ensure_red(c, (nb + 1) * 7 * (long long)sizeof(T)))) return rc;
    T* part = (T*)c->red;
    HIPCHK(Stages<T>::courants(c->g, ox, oy, oz, h, level, dt * dxi, dt * dyi, dt * dzi, part, s));
    T m[7];
    HIPCHK(hipMemcpyAsync(m, part + 7 * nb, sizeof m, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int q = 0; q < 7; q++) out[q] = (double)m[q];
    return MPDATA_OK;
}

template int reduce_courants<double>(mpdata_ctx*, const double*, const double*,
                                      const double*, const double*, int, double, double,
                                      double, double, double*, hipStream_t);
template int reduce_courants<float>(mpdata_ctx*, const float*, const float*, const float*,
                                     const float*, int, float, float, float, float, double*,
                                     hipStream_t);

}  // namespace mpdata

namespace {

template <typename T>
int reduce_norms(mpdata_ctx* c, const T* x, const T* x0, double* out, hipStream_t s) {
    if (any_null(x, out)) return fail(MPDATA_EINVAL, "null pointer");
    const long long nb = Stages<T>::reduce_blocks(c->g);
    int rc;
    if ((rc = ensure_red(c, nb * 4 * (long long)sizeof(double)))) return rc;
    double* part = (double*)c->red;
    HIPCHK(Stages<T>::norm_sums(c->g, x, x0, part, s));
    std::vector<double> hpart((size_t)nb * 4);
    HIPCHK(hipMemcpyAsync(hpart.data(), part, hpart.size() * sizeof(double),
                          hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int q = 0; q < 4; q++) {
        double acc = 0.0;
        for (long long b = 0; b < nb; b++) acc += hpart[4 * b + q];
        out[q] = acc;
    }
    return MPDATA_OK;
}

}  // namespace

// ============================================================================
extern "C" {

const char* mpdata_version(void) { return "mpdata_hip 0.1 (gfx950)"; }

const void* mpdata_addr(const void* p) { return p; }

int mpdata_is_device_ptr(const void* p) {
    if (!p) return 0;
    hipPointerAttribute_t attr;
    const bool dev = hipPointerGetAttributes(&attr, p) == hipSuccess &&
                     (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged);
    (void)hipGetLastError();
    return dev ? 1 : 0;
}

const char* mpdata_last_error(void) { return mpdata::last_error(); }

int mpdata_ctx_create(int np, int mp, int lp, int ih, int ibcx, int ibcy, int ibcz,
                      int real_bytes, mpdata_ctx** out) {
    if (!out) return fail(MPDATA_EINVAL, "null output pointer");
    *out = nullptr;
    if (np < 3 || mp < 3 || lp < 3) return fail(MPDATA_EINVAL, "np,mp,lp must be >= 3");
    if (ih < 1) return fail(MPDATA_EINVAL, "ih must be >= 1");
    if ((ibcx | ibcy | ibcz) & ~1) return fail(MPDATA_EINVAL, "ibc flags must be 0 or 1");
    if (real_bytes != 4 && real_bytes != 8) return fail(MPDATA_EINVAL, "real_bytes must be 4 or 8");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    mpdata_ctx* c = new mpdata_ctx();
    c->g = Geo{np, mp, lp, ih, ibcx, ibcy, ibcz};
    c->real_bytes = real_bytes;
    c->path = MPDATA_PATH_AUTO;
    c->device = dev;
    *out = c;
    return MPDATA_OK;
}

int mpdata_ctx_destroy(mpdata_ctx* ctx) {
    if (!ctx) return MPDATA_OK;
    for (void* p : ctx->owned) (void)hipFree(p);
    if (ctx->red) (void)hipFree(ctx->red);
    delete ctx;
    return MPDATA_OK;
}

int mpdata_ctx_set_path(mpdata_ctx* ctx, int path) {
    if (!ctx) return fail(MPDATA_EINVAL, "null context");
    if (path < MPDATA_PATH_AUTO || path > MPDATA_PATH_FAST) return fail(MPDATA_EINVAL, "path %d", path);
    ctx->path = path;
    return MPDATA_OK;
}

long long mpdata_ctx_scratch_bytes(const mpdata_ctx* ctx) { return ctx ? ctx->held : 0; }

#define DEFINE_REAL_API(T, SFX, RB)                                                              \
    int mpdata_upwind3d_gpubc##SFX(mpdata_ctx* ctx, const T* u1, const T* u2, const T* u3,       \
                                   T* x_in, T* xant, const T* xforc_impl, const T* xforc_expl,   \
                                   const T* rhr, const T* h, const T* bcx, const T* bcy,         \
                                   int ipoles0, int ibcx0, int ibcy0, int ibcz0, int np, int mp, \
                                   int lp, int ih, T dt, void* stream) {                         \
        (void)xforc_impl; (void)xforc_expl; (void)dt;                                            \
        int rc;                                                                                  \
        if ((rc = check_ctx(ctx, RB))) return rc;                                                \
        if ((rc = check_dims(ctx, ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih))) return rc;     \
        if (any_null(u1, u2, u3, x_in, xant, rhr, h)) return fail(MPDATA_EINVAL, "null field"); \
        return upwind_impl<T>(ctx, u1, u2, u3, x_in, xant, rhr, h, bcx, bcy,                     \
                              (hipStream_t)stream);                                              \
    }                                                                                            \
    static int antidiff##SFX(int variant, mpdata_ctx* ctx, const T* u1, const T* u2,             \
                             const T* u3, T* xant, T* x_in, const T* rhr, const T* h,            \
                             int ipoles0, int ibcx0, int ibcy0, int ibcz0, int np, int mp,       \
                             int lp, int ih, void* stream) {                                     \
        int rc;                                                                                  \
        if ((rc = check_ctx(ctx, RB))) return rc;                                                \
        if ((rc = check_dims(ctx, ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih))) return rc;     \
        if (any_null(u1, u2, u3, x_in, xant, rhr, h)) return fail(MPDATA_EINVAL, "null field"); \
        return stage_pass<T>(ctx, variant, 1, u1, u2, u3, xant, x_in, x_in, rhr, h, nullptr,     \
                             nullptr, nullptr, (hipStream_t)stream);                             \
    }                                                                                            \
    int mpdata_antidiff3d_gauge_gpubc##SFX(mpdata_ctx* ctx, int lupdatemulti, const T* u1,       \
                                           const T* u2, const T* u3, T* xant, T* x_in,           \
                                           const T* rhr, const T* h, int iflip, int ipoles0,     \
                                           int ibcx0, int ibcy0, int ibcz0, int np, int mp,      \
                                           int lp, int ih, void* stream) {                       \
        (void)lupdatemulti; (void)iflip;                                                         \
        return antidiff##SFX(MPDATA_GAUGE, ctx, u1, u2, u3, xant, x_in, rhr, h, ipoles0, ibcx0,  \
                             ibcy0, ibcz0, np, mp, lp, ih, stream);                              \
    }                                                                                            \
    int mpdata_antidiff3d_standard_gpubc##SFX(mpdata_ctx* ctx, int lupdatemulti, const T* u1,    \
                                              const T* u2, const T* u3, T* xant, T* x_in,        \
                                              const T* rhr, const T* h, int iflip, int ipoles0,  \
                                              int ibcx0, int ibcy0, int ibcz0, int np, int mp,   \
                                              int lp, int ih, void* stream) {                    \
        (void)lupdatemulti; (void)iflip;                                                         \
        return antidiff##SFX(MPDATA_STANDARD, ctx, u1, u2, u3, xant, x_in, rhr, h, ipoles0,      \
                             ibcx0, ibcy0, ibcz0, np, mp, lp, ih, stream);                       \
    }                                                                                            \
    int mpdata_advec_dwarf##SFX(mpdata_ctx* ctx, int legacyoptmode, int t_adv,                   \
                                int lupdatemulti, int lvertsplit, int np, int mp, int lp,        \
                                int ih, int ipoles, int ibcx, int ibcy, int ibcz, T* x,          \
                                const T* x_lr, const T* x_bt, const T* uadv, const T* vadv,      \
                                const T* wadv, const T* rhr, const T* rhoadv, void* stream) {    \
        (void)lupdatemulti;                                                                      \
        int rc;                                                                                  \
        if ((rc = check_ctx(ctx, RB))) return rc;                                                \
        if ((rc = check_dims(ctx, ipoles, ibcx, ibcy, ibcz, np, mp, lp, ih))) return rc;         \
        if (lvertsplit) return fail(MPDATA_EUNSUPPORTED, "lvertsplit (commented out in advec.f90)"); \
        int variant;                                                                             \
        if (legacyoptmode == 6)                                                                  \
            variant = MPDATA_UPWIND;                                                             \
        else if (legacyoptmode == 4 && (t_adv == 1 || t_adv == 2))                               \
            variant = t_adv == 1 ? MPDATA_GAUGE : MPDATA_STANDARD;                               \
        else                                                                                     \
            return fail(MPDATA_EINVAL,                                                           \
                        "legacyoptmode %d / t_adv %d: advec.f90 accepts 4 with t_adv 1|2, or 6", \
                        legacyoptmode, t_adv);                                                   \
        /* the two-step driver's timer, advec.f90:183-240 (none in mode 6) */                   \
        const int tid = legacyoptmode == 6 ? 0 : (t_adv == 1 ? 55 : 57);                        \
        if (tid) tt_beg(tid, (hipStream_t)stream);                                               \
        rc = step_impl<T>(ctx, variant, 2, 1, x, uadv, vadv, wadv, rhr, rhoadv, x_lr, x_bt,      \
                          (hipStream_t)stream);                                                  \
        if (tid) tt_end(tid, (hipStream_t)stream);                                               \
        return rc;                                                                               \
    }                                                                                            \
    int mpdata_step##SFX(mpdata_ctx* ctx, int variant, int iters, int nonos, T* x, const T* u1,  \
                         const T* u2, const T* u3, const T* rhr, const T* h, const T* bcx,       \
                         const T* bcy, void* stream) {                                           \
        int rc;                                                                                  \
        if ((rc = check_ctx(ctx, RB))) return rc;                                                \
        return step_impl<T>(ctx, variant, iters, nonos, x, u1, u2, u3, rhr, h, bcx, bcy,         \
                            (hipStream_t)stream);                                                \
    }                                                                                            \
    int mpdata_c_timeloop##SFX(mpdata_ctx* ctx, int variant, int iters, int nonos, int nt, T* x, \
                               const T* u1, const T* u2, const T* u3, const T* rhr, const T* h,  \
                               const T* bcx, const T* bcy, void* stream) {                       \
        int rc;                                                                                  \
        if ((rc = check_ctx(ctx, RB))) return rc;                                                \
        return timeloop_impl<T>(ctx, variant, iters, nonos, nt, x, u1, u2, u3, rhr, h, bcx, bcy, \
                                (hipStream_t)stream);                                            \
    }                                                                                            \
    int mpdata_step_oop##SFX(mpdata_ctx* ctx, int variant, int iters, int nonos, const T* x_in,  \
                             T* x_out, const T* u1, const T* u2, const T* u3, const T* rhr,      \
                             const T* h, const T* bcx, const T* bcy, void* stream) {             \
        int rc;                                                                                  \
        if ((rc = check_ctx(ctx, RB))) return rc;                                                \
        return step_oop_impl<T>(ctx, variant, iters, nonos, x_in, x_out, u1, u2, u3, rhr, h,     \
                                bcx, bcy, (hipStream_t)stream);                                  \
    }                                                                                            \
    int mpdata_velprd_traj0##SFX(mpdata_ctx* ctx, T* ox, T* oy, T* oz, T* u1, T* u2, T* u3,      \
                                 const T* h, T dxi, T dyi, T dzi, T dt, T dt_ratio,              \
                                 void* stream) {                                                 \
        int rc;                                                                                  \
        if ((rc = check_ctx(ctx, RB))) return rc;                                                \
        if (any_null(ox, oy, oz, u1, u2, u3, h)) return fail(MPDATA_EINVAL, "null field");      \
        const Geo& g = ctx->g;                                                                   \
        const long long n = len_of<SCAL>(g);                                                     \
        hipStream_t s = (hipStream_t)stream;                                                     \
        HIPCHK(Stages<T>::velprd_extrap(g, ox, oy, oz, h, dt * dxi, dt * dyi, dt * dzi, dt_ratio, \
                                        s));                                                     \
        HIPCHK(Stages<T>::a_to_c(g, ox + n, oy + n, oz + n, u1, u2, u3, s));                     \
        return MPDATA_OK;                                                                        \
    }                                                                                            \
    int mpdata_velprd_c##SFX(mpdata_ctx* ctx, const T* ox, const T* oy, const T* oz,             \
                             const T* oxp, const T* oyp, const T* ozp, const T* h, T dxi, T dyi, \
                             T dzi, T dt, T dt_ratio, T* u1, T* u2, T* u3, void* stream) {      \
        int rc;                                                                                  \
        if ((rc = check_ctx(ctx, RB))) return rc;                                                \
        if (any_null(ox, oy, oz, oxp, oyp, ozp, h, u1, u2, u3))                                  \
            return fail(MPDATA_EINVAL, "null field");                                            \
        const Geo& g = ctx->g;                                                                   \
        if (g.px || g.py || g.pz)                                                                \
            return fail(MPDATA_EUNSUPPORTED, "velprd_c: open boundaries only");                 \
        HIPCHK(Stages<T>::predict_ac(g, ox, oy, oz, oxp, oyp, ozp, h, dt * dxi, dt * dyi,        \
                                     dt * dzi, dt_ratio, u1, u2, u3, (hipStream_t)stream));      \
        return MPDATA_OK;                                                                        \
    }                                                                                            \
    int mpdata_velprdA_to_C##SFX(mpdata_ctx* ctx, T* ox, T* oy, T* oz, T* u1, T* u2, T* u3,      \
                                 void* stream) {                                                 \
        int rc;                                                                                  \
        if ((rc = check_ctx(ctx, RB))) return rc;                                                \
        if (any_null(ox, oy, oz, u1, u2, u3)) return fail(MPDATA_EINVAL, "null field");         \
        HIPCHK(Stages<T>::a_to_c(ctx->g, ox, oy, oz, u1, u2, u3, (hipStream_t)stream));          \
        return MPDATA_OK;                                                                        \
    }                                                                                            \
    int mpdata_velprd_rotate##SFX(mpdata_ctx* ctx, T* u, T* v, T* w, T* ox, T* oy, T* oz,        \
                                  const T* h, T dt_ratio, void* stream) {                        \
        int rc;                                                                                  \
        if ((rc = check_ctx(ctx, RB))) return rc;                                                \
        if (any_null(u, v, w, ox, oy, oz, h)) return fail(MPDATA_EINVAL, "null field");         \
        HIPCHK(Stages<T>::velprd_rotate(ctx->g, u, v, w, ox, oy, oz, h, dt_ratio,                \
                                        (hipStream_t)stream));                                   \
        return MPDATA_OK;                                                                        \
    }                                                                                            \
    int mpdata_compute_courants##SFX(mpdata_ctx* ctx, const T* ox, const T* oy, const T* oz,     \
                                     const T* h, int ntimelevel, T dxi, T dyi, T dzi, T dt,      \
                                     double* out, void* stream) {                                \
        int rc;                                                                                  \
        if ((rc = check_ctx(ctx, RB))) return rc;                                                \
        return reduce_courants<T>(ctx, ox, oy, oz, h, ntimelevel, dxi, dyi, dzi, dt, out,        \
                                  (hipStream_t)stream);                                          \
    }                                                                                            \
    int mpdata_norm_sums##SFX(mpdata_ctx* ctx, const T* x, const T* x0, double* out,             \
                              void* stream) {                                                    \
        int rc;                                                                                  \
        if ((rc = check_ctx(ctx, RB))) return rc;                                                \
        return reduce_norms<T>(ctx, x, x0, out, (hipStream_t)stream);                            \
    }

DEFINE_REAL_API(double, _f64, 8)
DEFINE_REAL_API(float, _f32, 4)

int upwind_c_timeloop(void* xtracer, void* uadv, void* vadv, void* wadv, void* rhr, void* h,
                      int np, int mp, int lp, int ih, int nt) {
    if (any_null(xtracer, uadv, vadv, wadv, rhr, h)) return fail(MPDATA_EINVAL, "null pointer");
    if (nt < 0) return fail(MPDATA_EINVAL, "nt %d", nt);
    mpdata_ctx* c = nullptr;
    int rc = mpdata_ctx_create(np, mp, lp, ih, 0, 0, 0, 8, &c);
    if (rc) return rc;
    const Geo g = c->g;
    const size_t sizes[6] = {(size_t)len_of<SCAL>(g), (size_t)len_of<FACEX>(g),
                             (size_t)len_of<FACEY>(g), (size_t)len_of<FACEZ>(g),
                             (size_t)len_of<SCAL>(g), (size_t)len_of<SCAL>(g)};
    void* user[6] = {xtracer, uadv, vadv, wadv, rhr, h};
    void* dev[6] = {nullptr};
    bool staged[6] = {false};
    double* bc = nullptr;
    const size_t nbc = 2 * (size_t)mp * lp + 2 * (size_t)np * lp;
    hipError_t e = hipSuccess;
    for (int a = 0; a < 6 && e == hipSuccess; a++) {
        hipPointerAttribute_t attr;
        bool on_device = hipPointerGetAttributes(&attr, user[a]) == hipSuccess &&
                         (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged);
        (void)hipGetLastError();
        if (on_device) {
            dev[a] = user[a];
            continue;
        }
        e = hipMalloc(&dev[a], sizes[a] * sizeof(double));
        staged[a] = e == hipSuccess;
        if (e == hipSuccess)
            e = hipMemcpy(dev[a], user[a], sizes[a] * sizeof(double), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMalloc(&bc, nbc * sizeof(double));
    if (e == hipSuccess) e = hipMemset(bc, 0, nbc * sizeof(double));
    if (e != hipSuccess) {
        rc = hip_fail(e, "upwind_c_timeloop staging");
    } else {
        for (int it = 0; it < nt && rc == MPDATA_OK; it++)
            rc = step_impl<double>(c, MPDATA_UPWIND, 2, 1, (double*)dev[0], (double*)dev[1],
                                   (double*)dev[2], (double*)dev[3], (double*)dev[4],
                                   (double*)dev[5], bc, bc + 2 * (size_t)mp * lp, nullptr);
        if (rc == MPDATA_OK && staged[0]) {
            e = hipMemcpy(xtracer, dev[0], sizes[0] * sizeof(double), hipMemcpyDeviceToHost);
            if (e != hipSuccess) rc = hip_fail(e, "upwind_c_timeloop copy-out");
        }
        if (rc == MPDATA_OK) {
            e = hipDeviceSynchronize();
            if (e != hipSuccess) rc = hip_fail(e, "upwind_c_timeloop");
        }
    }
    for (int a = 0; a < 6; a++)
        if (staged[a]) (void)hipFree(dev[a]);
    if (bc) (void)hipFree(bc);
    mpdata_ctx_destroy(c);
    return rc;
}

// ---- tangent-linear step ----------------------------------------------------
}  // extern "C"

namespace {

using D64 = Dual<double>;

__global__ void k_pack_dual(const double* __restrict__ v, const double* __restrict__ d,
                            D64* __restrict__ out, long long n) {
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
         q += (long long)gridDim.x * blockDim.x)
        out[q] = D64(v[q], d ? d[q] : 0.0);
}

__global__ void k_unpack_dual(const D64* __restrict__ in, double* __restrict__ v,
                              double* __restrict__ d, long long n) {
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
         q += (long long)gridDim.x * blockDim.x) {
        const D64 a = in[q];
        v[q] = a.v;
        d[q] = a.d;
    }
}

int grid_of(long long n) { return (int)std::min<long long>(2048, (n + 255) / 256); }

int tl_step_impl(mpdata_ctx* c, int variant, int iters, int nonos, double* x, double* dx,
                 const double* u1, const double* u2, const double* u3, const double* du1,
                 const double* du2, const double* du3, const double* rhr, const double* h,
                 const double* bcx, const double* bcy, hipStream_t s) {
    int rc;
    if ((rc = check_step_opts(variant, iters, nonos))) return rc;
    if ((rc = check_step_args<double>(c, x, u1, u2, u3, rhr, h, bcx, bcy))) return rc;
    if (!dx) return fail(MPDATA_EINVAL, "null tangent dx");
    Scratch& S = c->dual;
    const Geo& g = c->g;
    const int kinds[6] = {SCAL, FACEX, FACEY, FACEZ, SCAL, SCAL};
    for (int q = 0; q < 6; q++)
        if ((rc = ensure(c, &S.in[q], kinds[q], sizeof(D64)))) return rc;
    // inflow values: (2, lp, mp) and (2, lp, np) arrays, sized as x faces to fit
    if ((rc = ensure(c, &S.in[6], FACEX, sizeof(D64))) || (rc = ensure(c, &S.in[7], FACEX, sizeof(D64))))
        return rc;
    const double* vals[6] = {x, u1, u2, u3, rhr, h};
    const double* tans[6] = {dx, du1, du2, du3, nullptr, nullptr};
    for (int q = 0; q < 6; q++) {
        const long long n = field_len(g, kinds[q]);
        k_pack_dual<<<grid_of(n), 256, 0, s>>>(vals[q], tans[q], (D64*)S.in[q], n);
    }
    if (bcx) {
        const long long n = 2ll * g.mp * g.lp;
        k_pack_dual<<<grid_of(n), 256, 0, s>>>(bcx, nullptr, (D64*)S.in[6], n);
    }
    if (bcy) {
        const long long n = 2ll * g.np * g.lp;
        k_pack_dual<<<grid_of(n), 256, 0, s>>>(bcy, nullptr, (D64*)S.in[7], n);
    }
    HIPCHK(hipGetLastError());
    D64* X = (D64*)S.in[0];
    rc = step_stages<D64>(c, variant, iters, nonos, X, (const D64*)S.in[1], (const D64*)S.in[2],
                          (const D64*)S.in[3], (const D64*)S.in[4], (const D64*)S.in[5],
                          bcx ? (const D64*)S.in[6] : nullptr, bcy ? (const D64*)S.in[7] : nullptr, s);
    if (rc) return rc;
    const long long n = field_len(g, SCAL);
    k_unpack_dual<<<grid_of(n), 256, 0, s>>>(X, x, dx, n);
    HIPCHK(hipGetLastError());
    return MPDATA_OK;
}

}  // namespace

extern "C" {

int mpdata_tl_step_f64(mpdata_ctx* ctx, int variant, int iters, int nonos, double* x, double* dx,
                       const double* u1, const double* u2, const double* u3, const double* du1,
                       const double* du2, const double* du3, const double* rhr, const double* h,
                       const double* bcx, const double* bcy, void* stream) {
    int rc = check_ctx(ctx, 8);
    if (rc) return rc;
    return tl_step_impl(ctx, variant, iters, nonos, x, dx, u1, u2, u3, du1, du2, du3, rhr, h, bcx,
                        bcy, (hipStream_t)stream);
}

int mpdata_dev_alloc(long long bytes, void** ptr) {
    if (!ptr) return fail(MPDATA_EINVAL, "null output pointer");
    *ptr = nullptr;
    if (bytes <= 0) return fail(MPDATA_EINVAL, "bytes must be positive (got %lld)", bytes);
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, (size_t)bytes);
    if (e != hipSuccess) return fail(MPDATA_ENOMEM, "hipMalloc(%lld): %s", bytes, hipGetErrorString(e));
    e = hipMemset(p, 0, (size_t)bytes);
    if (e != hipSuccess) {
        (void)hipFree(p);
        return hip_fail(e, "hipMemset");
    }
    *ptr = p;
    return MPDATA_OK;
}

int mpdata_dev_free(void* ptr) {
    if (!ptr) return MPDATA_OK;
    HIPCHK(hipFree(ptr));
    return MPDATA_OK;
}

int mpdata_copy(void* dst, const void* src, long long bytes, int kind) {
    static const hipMemcpyKind kinds[3] = {hipMemcpyHostToDevice, hipMemcpyDeviceToHost,
                                           hipMemcpyDeviceToDevice};
    if (kind < 0 || kind > 2) return fail(MPDATA_EINVAL, "copy kind %d (0 h2d, 1 d2h, 2 d2d)", kind);
    if (bytes < 0) return fail(MPDATA_EINVAL, "negative byte count");
    if (bytes == 0) return MPDATA_OK;
    if (!dst || !src) return fail(MPDATA_EINVAL, "null pointer");
    HIPCHK(hipMemcpy(dst, src, (size_t)bytes, kinds[kind]));
    return MPDATA_OK;
}

int mpdata_sync(void* stream) {
    if (stream) HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    else HIPCHK(hipDeviceSynchronize());
    return MPDATA_OK;
}

}  // extern "C"
