// Ensembles: many MPDATA domains of different shapes (tracers, ensemble
// members) advanced together by the batched fused kernel, with an autotuner
// for its tile configuration and k-chunk depth and a HIP-graph time loop
// (C ABI in include/mpdata_hip.h).
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <vector>

#include "../../include/mpdata_hip.h"
#include "mpdata_fused.hpp"
#include "mpdata_status.hpp"

using namespace mpdata;

struct mpdata_ens {
    int n = 0, ih = 1, real_bytes = 8, device = 0;
    std::vector<Geo> g;
    // member fields (device pointers); x alternates between xa and xb
    std::vector<const void*> u1, u2, u3, rhr, h, bcx, bcy;
    std::vector<void*> xa, xb;
    bool have_fields = false;
    int cfg = 0, clen = 64;
    int fast = 0;  // the fast build of the fused step (MPDATA_PATH_FAST)
    int nitem = 0;
    long long max_halo = 0;
    void* table[2] = {nullptr, nullptr};  // parity 0: xa -> xb, 1: xb -> xa
    bool tables_ok = false;
    // cached graph of an nsteps-step loop
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_steps = -1;
};

namespace {

// Autotuned (cfg, clen) per ensemble signature (shapes + real size), kept for
// the life of the process.
std::mutex g_tune_mu;
std::map<std::vector<int>, std::pair<int, int>> g_tuned;

std::vector<int> signature(const mpdata_ens* e) {
    std::vector<int> s{e->real_bytes, e->ih, e->n};
    for (const Geo& g : e->g) {
        s.push_back(g.np);
        s.push_back(g.mp);
        s.push_back(g.lp);
    }
    return s;
}

void drop_graph(mpdata_ens* e) {
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
    e->graph_steps = -1;
}

template <typename T>
int pack_table(mpdata_ens* e, int cfg, int clen, int parity, std::vector<char>& host) {
    std::vector<BatchMember<T>> m(e->n);
    for (int q = 0; q < e->n; q++) {
        BatchMember<T>& b = m[q];
        b.g = e->g[q];
        b.x = (const T*)(parity ? e->xb[q] : e->xa[q]);
        b.xo = (T*)(parity ? e->xa[q] : e->xb[q]);
        b.u1 = (const T*)e->u1[q];
        b.u2 = (const T*)e->u2[q];
        b.u3 = (const T*)e->u3[q];
        b.rhr = (const T*)e->rhr[q];
        b.h = (const T*)e->h[q];
        b.bcx = (const T*)e->bcx[q];
        b.bcy = (const T*)e->bcy[q];
    }
    host.assign(FusedBatch<T>::table_bytes(e->n), 0);
    return FusedBatch<T>::pack(m.data(), e->n, cfg, clen, host.data());
}

int pack_table_any(mpdata_ens* e, int cfg, int clen, int parity, std::vector<char>& host) {
    return e->real_bytes == 8 ? pack_table<double>(e, cfg, clen, parity, host)
                              : pack_table<float>(e, cfg, clen, parity, host);
}

// Packs and uploads both parity tables for (cfg, clen).
int build_tables(mpdata_ens* e, int cfg, int clen) {
    // no step or graph replay may still read the old tables, and a graph
    // exec is destroyed only once no replay of it runs
    HIPCHK(hipDeviceSynchronize());
    drop_graph(e);
    e->tables_ok = false;
    std::vector<char> host;
    int nitem = -1;
    for (int parity = 0; parity < 2; parity++) {
        nitem = pack_table_any(e, cfg, clen, parity, host);
        if (nitem < 0) return fail(MPDATA_EINVAL, "ensemble: cannot pack configuration %d", cfg);
        HIPCHK(hipMemcpy(e->table[parity], host.data(), host.size(), hipMemcpyHostToDevice));
    }
    e->cfg = cfg;
    e->clen = clen;
    e->nitem = nitem;
    e->tables_ok = true;
    return MPDATA_OK;
}

hipError_t launch_step(const mpdata_ens* e, int parity, hipStream_t s) {
    if (e->fast)
        return e->real_bytes == 8
                   ? FusedBatchFast<double>::step(e->table[parity], e->n, e->nitem, e->cfg,
                                                  e->max_halo, s)
                   : FusedBatchFast<float>::step(e->table[parity], e->n, e->nitem, e->cfg,
                                                 e->max_halo, s);
    return e->real_bytes == 8
               ? FusedBatch<double>::step(e->table[parity], e->n, e->nitem, e->cfg, e->max_halo, s)
               : FusedBatch<float>::step(e->table[parity], e->n, e->nitem, e->cfg, e->max_halo, s);
}

int check_ens(const mpdata_ens* e, int real_bytes) {
    if (!e) return fail(MPDATA_EINVAL, "null ensemble");
    if (real_bytes && e->real_bytes != real_bytes)
        return fail(MPDATA_EINVAL, "ensemble holds %d-byte reals, call needs %d", e->real_bytes,
                    real_bytes);
    int dev = -1;
    if (hipGetDevice(&dev) == hipSuccess && dev != e->device)
        return fail(MPDATA_EINVAL, "ensemble was created on device %d, current device is %d",
                    e->device, dev);
    return MPDATA_OK;
}

int check_ready(const mpdata_ens* e) {
    if (!e->have_fields) return fail(MPDATA_EINVAL, "ensemble: fields not set");
    if (!e->tables_ok) return fail(MPDATA_EINVAL, "ensemble: no packed configuration");
    return MPDATA_OK;
}

template <typename T>
int set_fields(mpdata_ens* e, T* const* xa, T* const* xb, const T* const* u1, const T* const* u2,
               const T* const* u3, const T* const* rhr, const T* const* h, const T* const* bcx,
               const T* const* bcy) {
    if (!xa || !xb || !u1 || !u2 || !u3 || !rhr || !h || !bcx || !bcy)
        return fail(MPDATA_EINVAL, "ensemble: null pointer array");
    for (int q = 0; q < e->n; q++) {
        if (!xa[q] || !xb[q] || !u1[q] || !u2[q] || !u3[q] || !rhr[q] || !h[q] || !bcx[q] ||
            !bcy[q])
            return fail(MPDATA_EINVAL, "ensemble: null field of member %d", q);
        const Geo& g = e->g[q];
        const size_t bytes = (size_t)len_of<SCAL>(g) * sizeof(T);
        const char *a = (const char*)xa[q], *b = (const char*)xb[q];
        if (a < b + bytes && b < a + bytes)
            return fail(MPDATA_EINVAL, "ensemble: x buffers of member %d overlap", q);
    }
    e->xa.assign(xa, xa + e->n);
    e->xb.assign(xb, xb + e->n);
    e->u1.assign(u1, u1 + e->n);
    e->u2.assign(u2, u2 + e->n);
    e->u3.assign(u3, u3 + e->n);
    e->rhr.assign(rhr, rhr + e->n);
    e->h.assign(h, h + e->n);
    e->bcx.assign(bcx, bcx + e->n);
    e->bcy.assign(bcy, bcy + e->n);
    e->have_fields = true;
    return build_tables(e, e->cfg, e->clen);
}

}  // namespace

extern "C" {

int mpdata_ens_create(int nmember, const int* np, const int* mp, const int* lp, int ih,
                      int real_bytes, mpdata_ens** out) {
    if (!out) return fail(MPDATA_EINVAL, "null output pointer");
    *out = nullptr;
    if (nmember < 1 || !np || !mp || !lp) return fail(MPDATA_EINVAL, "ensemble: no members");
    if (ih != 1) return fail(MPDATA_EUNSUPPORTED, "ensemble: ih must be 1 (got %d)", ih);
    if (real_bytes != 4 && real_bytes != 8) return fail(MPDATA_EINVAL, "real_bytes must be 4 or 8");
    std::vector<Geo> geos(nmember);
    long long max_halo = 0;
    for (int q = 0; q < nmember; q++) {
        if (np[q] < 3 || mp[q] < 3 || lp[q] < 3)
            return fail(MPDATA_EINVAL, "ensemble member %d: np,mp,lp must be >= 3", q);
        geos[q] = Geo{np[q], mp[q], lp[q], ih, 0, 0, 0};
        if (!Fused<double>::supports(geos[q], MPDATA_GAUGE, 2, 1))
            return fail(MPDATA_EUNSUPPORTED, "ensemble member %d is too large for 32-bit offsets", q);
        const long long nx = np[q] + 2 * ih, ny = mp[q] + 2 * ih;
        const long long halo = 2ll * ih * nx * ny + 2ll * ih * nx * lp[q] + 2ll * ih * mp[q] * lp[q];
        if (halo > max_halo) max_halo = halo;
    }
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    mpdata_ens* e = new mpdata_ens();
    e->n = nmember;
    e->ih = ih;
    e->real_bytes = real_bytes;
    e->device = dev;
    e->g = geos;
    e->max_halo = max_halo;
    const size_t tb = real_bytes == 8 ? FusedBatch<double>::table_bytes(nmember)
                                      : FusedBatch<float>::table_bytes(nmember);
    for (int p = 0; p < 2; p++) {
        hipError_t err = hipMalloc(&e->table[p], tb);
        if (err != hipSuccess) {
            (void)hipGetLastError();
            if (e->table[0]) (void)hipFree(e->table[0]);
            delete e;
            return fail(MPDATA_ENOMEM, "ensemble: hipMalloc(%zu) failed", tb);
        }
    }
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(signature(e));
        if (it != g_tuned.end()) {
            e->cfg = it->second.first;
            e->clen = it->second.second;
        }
    }
    *out = e;
    return MPDATA_OK;
}

int mpdata_ens_destroy(mpdata_ens* e) {
    if (!e) return MPDATA_OK;
    (void)hipDeviceSynchronize();
    drop_graph(e);
    if (e->cap) (void)hipStreamDestroy(e->cap);
    for (void* t : e->table)
        if (t) (void)hipFree(t);
    delete e;
    return MPDATA_OK;
}

int mpdata_ens_set_fields_f64(mpdata_ens* e, double* const* xa, double* const* xb,
                              const double* const* u1, const double* const* u2,
                              const double* const* u3, const double* const* rhr,
                              const double* const* h, const double* const* bcx,
                              const double* const* bcy) {
    int rc = check_ens(e, 8);
    if (rc) return rc;
    return set_fields<double>(e, xa, xb, u1, u2, u3, rhr, h, bcx, bcy);
}

int mpdata_ens_set_fields_f32(mpdata_ens* e, float* const* xa, float* const* xb,
                              const float* const* u1, const float* const* u2,
                              const float* const* u3, const float* const* rhr,
                              const float* const* h, const float* const* bcx,
                              const float* const* bcy) {
    int rc = check_ens(e, 4);
    if (rc) return rc;
    return set_fields<float>(e, xa, xb, u1, u2, u3, rhr, h, bcx, bcy);
}

int mpdata_ens_nconfigs(void) { return FusedBatch<double>::nconfigs(); }

const char* mpdata_ens_config_name(int cfg) {
    if (cfg < 0 || cfg >= FusedBatch<double>::nconfigs()) return "";
    return FusedBatch<double>::config_name(cfg);
}

int mpdata_ens_set_tiling(mpdata_ens* e, int cfg, int clen) {
    int rc = check_ens(e, 0);
    if (rc) return rc;
    if (cfg < 0 || cfg >= FusedBatch<double>::nconfigs())
        return fail(MPDATA_EINVAL, "ensemble: tile configuration %d (0..%d)", cfg,
                    FusedBatch<double>::nconfigs() - 1);
    if (clen < 1) return fail(MPDATA_EINVAL, "ensemble: k-chunk length %d", clen);
    if (!e->have_fields) {
        e->cfg = cfg;
        e->clen = clen;
        return MPDATA_OK;
    }
    return build_tables(e, cfg, clen);
}

int mpdata_ens_set_path(mpdata_ens* e, int path) {
    int rc = check_ens(e, 0);
    if (rc) return rc;
    if (path != MPDATA_PATH_AUTO && path != MPDATA_PATH_FUSED && path != MPDATA_PATH_FAST)
        return fail(MPDATA_EUNSUPPORTED,
                    "ensemble: path %d (the batch runs the fused step: AUTO, FUSED or FAST)", path);
    const int fast = path == MPDATA_PATH_FAST;
    if (fast != e->fast) {
        // no replay of a graph of the other build may still run
        HIPCHK(hipDeviceSynchronize());
        drop_graph(e);
        e->fast = fast;
    }
    return MPDATA_OK;
}

int mpdata_ens_info(const mpdata_ens* e, int* nitem, int* cfg, int* clen) {
    int rc = check_ens(e, 0);
    if (rc) return rc;
    if (nitem) *nitem = e->tables_ok ? e->nitem : 0;
    if (cfg) *cfg = e->cfg;
    if (clen) *clen = e->clen;
    return MPDATA_OK;
}

int mpdata_ens_step(mpdata_ens* e, int parity, void* stream) {
    int rc = check_ens(e, 0);
    if (rc || (rc = check_ready(e))) return rc;
    if (parity != 0 && parity != 1) return fail(MPDATA_EINVAL, "parity must be 0 or 1");
    HIPCHK(launch_step(e, parity, (hipStream_t)stream));
    return MPDATA_OK;
}

int mpdata_ens_run(mpdata_ens* e, int nsteps, int use_graph, void* stream) {
    int rc = check_ens(e, 0);
    if (rc || (rc = check_ready(e))) return rc;
    if (nsteps < 0) return fail(MPDATA_EINVAL, "nsteps must be >= 0");
    if (nsteps == 0) return MPDATA_OK;
    hipStream_t s = (hipStream_t)stream;
    if (!use_graph) {
        for (int k = 0; k < nsteps; k++) HIPCHK(launch_step(e, k & 1, s));
        return MPDATA_OK;
    }
    if (!e->exec || e->graph_steps != nsteps) {
        // the previous graph may still be replaying on any stream
        if (e->exec) HIPCHK(hipDeviceSynchronize());
        drop_graph(e);
        if (!e->cap) HIPCHK(hipStreamCreateWithFlags(&e->cap, hipStreamNonBlocking));
        HIPCHK(hipStreamBeginCapture(e->cap, hipStreamCaptureModeRelaxed));
        hipError_t err = hipSuccess;
        for (int k = 0; k < nsteps && err == hipSuccess; k++) err = launch_step(e, k & 1, e->cap);
        hipGraph_t graph = nullptr;
        hipError_t end = hipStreamEndCapture(e->cap, &graph);
        if (err != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return hip_fail(err, "ensemble graph capture");
        }
        if (end != hipSuccess) return hip_fail(end, "hipStreamEndCapture");
        e->graph = graph;
        HIPCHK(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
        e->graph_steps = nsteps;
    }
    HIPCHK(hipGraphLaunch(e->exec, s));
    return MPDATA_OK;
}

int mpdata_ens_autotune(mpdata_ens* e, int trial_steps, void* stream, int* cfg_out, int* clen_out,
                        double* ms_out) {
    int rc = check_ens(e, 0);
    if (rc) return rc;
    if (!e->have_fields) return fail(MPDATA_EINVAL, "ensemble: fields not set");
    if (trial_steps < 1) return fail(MPDATA_EINVAL, "trial_steps must be >= 1");
    hipStream_t s = (hipStream_t)stream;
    static const int clens[] = {16, 32, 64, 128};
    int best_cfg = -1, best_clen = 0;
    float best_ms = 0.f;
    hipEvent_t t0, t1;
    HIPCHK(hipEventCreate(&t0));
    HIPCHK(hipEventCreate(&t1));
    int maxlp = 0;
    for (const Geo& g : e->g) maxlp = g.lp > maxlp ? g.lp : maxlp;
    for (int cfg = 0; cfg < FusedBatch<double>::nconfigs(); cfg++) {
        for (int clen : clens) {
            if (clen > 16 && clen / 2 >= maxlp) break;  // longer chunks change nothing
            if ((rc = build_tables(e, cfg, clen))) break;
            // parity 0 only: reads xa, overwrites xb
            hipError_t err = launch_step(e, 0, s);
            if (err == hipSuccess) err = hipEventRecord(t0, s);
            for (int k = 0; k < trial_steps && err == hipSuccess; k++) err = launch_step(e, 0, s);
            if (err == hipSuccess) err = hipEventRecord(t1, s);
            if (err == hipSuccess) err = hipEventSynchronize(t1);
            float ms = 0.f;
            if (err == hipSuccess) err = hipEventElapsedTime(&ms, t0, t1);
            if (err != hipSuccess) {
                rc = hip_fail(err, "ensemble autotune");
                break;
            }
            ms /= trial_steps;
            if (best_cfg < 0 || ms < best_ms) {
                best_cfg = cfg;
                best_clen = clen;
                best_ms = ms;
            }
        }
        if (rc) break;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (rc) return rc;
    if ((rc = build_tables(e, best_cfg, best_clen))) return rc;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[signature(e)] = {best_cfg, best_clen};
    }
    if (cfg_out) *cfg_out = best_cfg;
    if (clen_out) *clen_out = best_clen;
    if (ms_out) *ms_out = best_ms;
    return MPDATA_OK;
}

}  // extern "C"
