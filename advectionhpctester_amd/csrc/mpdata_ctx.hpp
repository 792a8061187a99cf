// The per-domain context behind the C ABI's opaque mpdata_ctx: geometry,
// execution path and device scratch (allocated on first use, zero-filled
// once, owned until mpdata_ctx_destroy).
#pragma once
#include <vector>

#include "mpdata_common.hpp"

namespace mpdata {

// One set of scratch fields; element size fixed by the set's use.
struct Scratch {
    void* xant = nullptr;
    void* psi2 = nullptr;
    void* xnext = nullptr;  // output buffer of the out-of-place fused step
    void* v[3] = {nullptr, nullptr, nullptr};
    void* cp = nullptr;
    void* cn = nullptr;
    void* fns[3] = {nullptr, nullptr, nullptr};
    void* f[3] = {nullptr, nullptr, nullptr};
    void* ul[3] = {nullptr, nullptr, nullptr};
    void* adv[3] = {nullptr, nullptr, nullptr};
    // packed inputs of the tangent-linear step (value, tangent pairs)
    void* in[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
};

// Forward intermediates of one recorded step (adjoint tape slot).
struct TapeSlot {
    void* xb = nullptr;     // x with the inflow halo (upwind input)
    void* x0 = nullptr;     // x with the edge-copied halo (limiter bounds)
    void* psi = nullptr;    // upwind result with the mirrored halo
    void* v[3] = {nullptr, nullptr, nullptr};  // pseudo-velocities
    void* cp = nullptr;
    void* cn = nullptr;
    int variant = 2;        // MPDATA_STANDARD or MPDATA_GAUGE
};

}  // namespace mpdata

struct mpdata_ctx {
    mpdata::Geo g;
    int real_bytes;
    int path;
    int device;
    mpdata::Scratch prim;   // scratch in the context's real type
    mpdata::Scratch dual;   // scratch of the tangent-linear step
    void* red = nullptr;    // reduction partials (device)
    long long red_bytes = 0;
    std::vector<mpdata::TapeSlot> tape;
    // adjoint work arrays (allocated zeroed; the reverse step leaves them
    // zeroed) and which of the two xant-adjoint buffers the next reverse
    // step accumulates into
    void* adj[24] = {};
    int adj_psi = 0;
    // same-grid multi-field step: output field per batched field, the
    // device descriptor table and the pointers it was packed for
    std::vector<void*> mf_out;
    void* mf_table = nullptr;
    size_t mf_table_bytes = 0;
    std::vector<const void*> mf_key;
    int mf_cfg = -1, mf_clen = 0, mf_nitem = 0;
    long long held = 0;     // bytes of device scratch
    std::vector<void*> owned;
};

namespace mpdata {

long long field_len(const Geo& g, int kind);
// Context of the right real size on the current device; shape and boundary
// flags of a reference-style argument list equal to the context's.
int check_ctx(const mpdata_ctx* c, int real_bytes);
int check_dims(const mpdata_ctx* c, int ipoles, int ibcx, int ibcy, int ibcz, int np, int mp,
               int lp, int ih);
// Device partials of the Courant maxima reduced on the host (synchronous):
// out = cr3d, crxy, crxz, cryz, crx, cry, crz (common/diagutils.f90:588-672).
template <typename T>
int reduce_courants(mpdata_ctx* c, const T* ox, const T* oy, const T* oz, const T* h, int level,
                    T dxi, T dyi, T dzi, T dt, double* out, hipStream_t s);

template <typename... P>
bool any_null(P... p) {
    const void* a[] = {(const void*)p...};
    for (const void* q : a)
        if (!q) return true;
    return false;
}
// Allocates *slot (field of `kind`, elem_bytes per element) unless set.
int ensure(mpdata_ctx* c, void** slot, int kind, int elem_bytes);

}  // namespace mpdata
