// Same-grid multi-field step: the fields of the reference's CE_advection
// (advection/src_drivers/advection_CEtimestep_twostep_GPUBC.f90:27-376)
// advanced with one velocity, one density and one metric.  CE_advection
// fills each field's inflow halo (putbcinxh, :160-177), runs the upwind
// pass of every field in one fused kernel (upwind_firstep_GPUBC_*) and then
// one antidiffusive pass per field, gauge for the fields it hands to
// mpdatm3d_antidiff (u, v, w, tht, pstr) and standard for those it hands to
// mpdata3d_antidiff (th and the moist fields) unless mpdata_alg_choice
// selects one variant for all (:145-155, :270-304); liner != 0 stops after
// the upwind pass (:330-341).  Per field that is exactly the reference's
// two-pass step, so each result equals mpdata_step on that field.
//
// Here every gauge field runs in one launch of the batched fused kernel in
// its shared-grid mode: the work items of one tile and k-chunk are
// consecutive for all fields, so one XCD marches them together and reads
// u1, u2, u3, h and rhr into its L2 once for all of them (on the fast build
// when the context is set to MPDATA_PATH_FAST, as mpdata_step_* is).
// Standard fields, and every field when the fused kernel does not apply,
// take per-field steps.
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/mpdata_hip.h"
#include "mpdata_ctx.hpp"
#include "mpdata_fused.hpp"
#include "mpdata_status.hpp"

using namespace mpdata;

namespace {

int variant_of(int alg_choice, int role) {
    if (alg_choice == 1) return MPDATA_GAUGE;
    if (alg_choice == 2) return MPDATA_STANDARD;
    return role ? MPDATA_GAUGE : MPDATA_STANDARD;  // 0 and > 2
}

// One field's step, in place when x_in == x_out.
template <typename T>
int step_one(mpdata_ctx* c, int variant, const T* x_in, T* x_out, const T* u1, const T* u2,
             const T* u3, const T* rhr, const T* h, const T* bcx, const T* bcy, void* s);
template <>
int step_one<double>(mpdata_ctx* c, int variant, const double* x_in, double* x_out,
                     const double* u1, const double* u2, const double* u3, const double* rhr,
                     const double* h, const double* bcx, const double* bcy, void* s) {
    if (x_in == x_out)
        return mpdata_step_f64(c, variant, 2, 1, x_out, u1, u2, u3, rhr, h, bcx, bcy, s);
    return mpdata_step_oop_f64(c, variant, 2, 1, x_in, x_out, u1, u2, u3, rhr, h, bcx, bcy, s);
}
template <>
int step_one<float>(mpdata_ctx* c, int variant, const float* x_in, float* x_out,
                    const float* u1, const float* u2, const float* u3, const float* rhr,
                    const float* h, const float* bcx, const float* bcy, void* s) {
    if (x_in == x_out)
        return mpdata_step_f32(c, variant, 2, 1, x_out, u1, u2, u3, rhr, h, bcx, bcy, s);
    return mpdata_step_oop_f32(c, variant, 2, 1, x_in, x_out, u1, u2, u3, rhr, h, bcx, bcy, s);
}

// Device table for the batched fields; re-packed and uploaded only when the
// fields, their outputs or the tiling differ from the previous call.
template <typename T>
int ensure_table(mpdata_ctx* c, const std::vector<BatchMember<T>>& m, int cfg, int clen,
                 hipStream_t s) {
    std::vector<const void*> key;
    key.reserve(m.size() * 11 + 2);
    for (const BatchMember<T>& b : m) {
        const void* p[] = {b.x, b.xo, b.u1, b.u2, b.u3, b.rhr, b.h, b.bcx, b.bcy};
        key.insert(key.end(), p, p + 9);
    }
    key.push_back((const void*)(long long)cfg);
    key.push_back((const void*)(long long)clen);
    if (c->mf_table && key == c->mf_key) return MPDATA_OK;
    const int n = (int)m.size();
    std::vector<char> host(FusedBatch<T>::table_bytes(n), 0);
    const int nitem = FusedBatch<T>::pack(m.data(), n, cfg, clen, host.data());
    if (nitem < 0) return fail(MPDATA_EINVAL, "multi-field step: cannot pack %d fields", n);
    // a step still in flight may read the old table
    HIPCHK(hipStreamSynchronize(s));
    if (c->mf_table_bytes < host.size()) {
        if (c->mf_table) {
            for (size_t i = 0; i < c->owned.size(); i++)
                if (c->owned[i] == c->mf_table) c->owned.erase(c->owned.begin() + i);
            (void)hipFree(c->mf_table);
            c->held -= (long long)c->mf_table_bytes;
            c->mf_table = nullptr;
            c->mf_table_bytes = 0;
        }
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, host.size());
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(MPDATA_ENOMEM, "hipMalloc(%zu bytes) for the field table failed", host.size());
        }
        c->mf_table = p;
        c->mf_table_bytes = host.size();
        c->owned.push_back(p);
        c->held += (long long)host.size();
    }
    HIPCHK(hipMemcpy(c->mf_table, host.data(), host.size(), hipMemcpyHostToDevice));
    c->mf_key = key;
    c->mf_cfg = cfg;
    c->mf_clen = clen;
    c->mf_nitem = nitem;
    return MPDATA_OK;
}

// In place (oop false): x_in and x_out are the same list.  Out of place:
// no output may overlap any field's input or another output.
template <typename T>
int ce_impl(mpdata_ctx* c, int alg_choice, int liner, int nfields, const int* role,
            const T* const* x_in, T* const* x_out, const T* const* bcx, const T* const* bcy,
            const T* u1, const T* u2, const T* u3, const T* rhr, const T* h, void* stream,
            bool oop) {
    if (!c) return fail(MPDATA_EINVAL, "null context");
    if (c->real_bytes != (int)sizeof(T))
        return fail(MPDATA_EINVAL, "context holds %d-byte reals, call needs %d", c->real_bytes,
                    (int)sizeof(T));
    if (nfields < 1 || nfields > MPDATA_MAX_FIELDS)
        return fail(MPDATA_EINVAL, "nfields %d outside 1..%d", nfields, MPDATA_MAX_FIELDS);
    if (alg_choice < 0) return fail(MPDATA_EINVAL, "mpdata_alg_choice %d", alg_choice);
    if (!role || !x_in || !x_out || !bcx || !bcy)
        return fail(MPDATA_EINVAL, "null field list");
    if (!u1 || !u2 || !u3 || !rhr || !h) return fail(MPDATA_EINVAL, "null field pointer");
    const size_t bytes = (size_t)len_of<SCAL>(c->g) * sizeof(T);
    auto overlap = [bytes](const void* p, const void* q) {
        const char *a = (const char*)p, *b = (const char*)q;
        return a < b + bytes && b < a + bytes;
    };
    for (int f = 0; f < nfields; f++) {
        if (role[f] != 0 && role[f] != 1) return fail(MPDATA_EINVAL, "role[%d] = %d", f, role[f]);
        if (!x_in[f] || !x_out[f]) return fail(MPDATA_EINVAL, "null field %d", f);
        for (int e = 0; e < nfields; e++) {
            if (e < f && overlap(x_out[f], x_out[e]))
                return fail(MPDATA_EINVAL, "fields %d and %d overlap", e, f);
            if (oop && overlap(x_out[f], x_in[e]))
                return fail(MPDATA_EINVAL, "output %d overlaps input %d", f, e);
        }
    }
    hipStream_t s = (hipStream_t)stream;
    if (liner) {
        for (int f = 0; f < nfields; f++) {
            int rc = step_one<T>(c, MPDATA_UPWIND, x_in[f], x_out[f], u1, u2, u3, rhr, h, bcx[f],
                                 bcy[f], stream);
            if (rc) return rc;
        }
        return MPDATA_OK;
    }
    // which fields go through the batched fused kernel
    const bool batch_ok = c->path != MPDATA_PATH_STAGES && c->g.ih == 1 &&
                          Fused<T>::supports(c->g, MPDATA_GAUGE, 2, 1);
    std::vector<int> batched;
    for (int f = 0; f < nfields; f++) {
        const int v = variant_of(alg_choice, role[f]);
        if (v == MPDATA_GAUGE && batch_ok) {
            batched.push_back(f);
            continue;
        }
        int rc = step_one<T>(c, v, x_in[f], x_out[f], u1, u2, u3, rhr, h, bcx[f], bcy[f], stream);
        if (rc) return rc;
    }
    if (batched.empty()) return MPDATA_OK;
    if (!bcx[batched[0]] && !c->g.px) return fail(MPDATA_EINVAL, "open boundaries need bcx/bcy");
    const int n = (int)batched.size();
    // fields updated in place go through an output field of the context
    if ((int)c->mf_out.size() < n) c->mf_out.resize(n, nullptr);
    std::vector<BatchMember<T>> m(n);
    for (int j = 0; j < n; j++) {
        const int f = batched[j];
        if (!bcx[f] || !bcy[f]) return fail(MPDATA_EINVAL, "field %d: open boundaries need bcx/bcy", f);
        const bool inplace = !oop;
        T* xo = x_out[f];
        if (inplace) {
            int rc = ensure(c, &c->mf_out[j], SCAL, sizeof(T));
            if (rc) return rc;
            xo = (T*)c->mf_out[j];
        }
        m[j] = BatchMember<T>{c->g, x_in[f], u1, u2, u3, rhr, h, bcx[f], bcy[f], xo};
    }
    int cfg = 0, clen = 64;
    FusedBatch<T>::choose_shared(c->g, n, cfg, clen);
    int rc = ensure_table<T>(c, m, cfg, clen, s);
    if (rc) return rc;
    const long long nx = c->g.np + 2, ny = c->g.mp + 2;
    const long long halo = 2 * nx * ny + 2 * nx * c->g.lp + 2ll * c->g.mp * c->g.lp;
    // the build of the fused step mpdata_step_* runs on this context
    if (c->path == MPDATA_PATH_FAST)
        HIPCHK(FusedBatchFast<T>::step(c->mf_table, n, c->mf_nitem, c->mf_cfg, halo, s, 1));
    else
        HIPCHK(FusedBatch<T>::step(c->mf_table, n, c->mf_nitem, c->mf_cfg, halo, s, 1));
    if (!oop)
        for (int j = 0; j < n; j++)
            HIPCHK(hipMemcpyAsync(x_out[batched[j]], c->mf_out[j], bytes, hipMemcpyDeviceToDevice, s));
    return MPDATA_OK;
}

}  // namespace

extern "C" {

int mpdata_ce_advection_f64(mpdata_ctx* ctx, int alg_choice, int liner, int nfields,
                            const int* role, double* const* x, const double* const* bcx,
                            const double* const* bcy, const double* u1, const double* u2,
                            const double* u3, const double* rhr, const double* h, void* stream) {
    return ce_impl<double>(ctx, alg_choice, liner, nfields, role, (const double* const*)x, x, bcx,
                           bcy, u1, u2, u3, rhr, h, stream, false);
}
int mpdata_ce_advection_f32(mpdata_ctx* ctx, int alg_choice, int liner, int nfields,
                            const int* role, float* const* x, const float* const* bcx,
                            const float* const* bcy, const float* u1, const float* u2,
                            const float* u3, const float* rhr, const float* h, void* stream) {
    return ce_impl<float>(ctx, alg_choice, liner, nfields, role, (const float* const*)x, x, bcx,
                          bcy, u1, u2, u3, rhr, h, stream, false);
}
int mpdata_ce_advection_oop_f64(mpdata_ctx* ctx, int alg_choice, int liner, int nfields,
                                const int* role, const double* const* x_in, double* const* x_out,
                                const double* const* bcx, const double* const* bcy,
                                const double* u1, const double* u2, const double* u3,
                                const double* rhr, const double* h, void* stream) {
    return ce_impl<double>(ctx, alg_choice, liner, nfields, role, x_in, x_out, bcx, bcy, u1, u2,
                           u3, rhr, h, stream, true);
}
int mpdata_ce_advection_oop_f32(mpdata_ctx* ctx, int alg_choice, int liner, int nfields,
                                const int* role, const float* const* x_in, float* const* x_out,
                                const float* const* bcx, const float* const* bcy,
                                const float* u1, const float* u2, const float* u3,
                                const float* rhr, const float* h, void* stream) {
    return ce_impl<float>(ctx, alg_choice, liner, nfields, role, x_in, x_out, bcx, bcy, u1, u2,
                          u3, rhr, h, stream, true);
}

}  // extern "C"
