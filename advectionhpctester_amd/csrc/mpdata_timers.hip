// Timers keyed by the reference's timer ids, measured with HIP events on the
// stream the timed work runs on: the analogue of its ttini / ttbeg / ttend /
// ttprt (common/mpi_parallel.f90:7088-7447, TIMERSCUDA), which time
// advec_dwarf (id 300 in the dwarf's loop, advection_dwarf_test_modular.f90:
// 211-218), the two-step driver (55 gauge / 57 standard, advec.f90:183-240),
// the upwind (16, upwind3d-gpubc.f90:87-200) and the antidiffusive passes
// (51 gauge / 53 standard).  Off until mpdata_ttini; the library records ids
// 55/57 around its step and 16, 51/53 around the per-stage passes.  A begin
// and an end record one event each; the elapsed times are summed when they
// are read (mpdata_ttget / mpdata_ttprt, which synchronize the events).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "../../include/mpdata_hip.h"
#include "mpdata_status.hpp"
#include "mpdata_timers.hpp"

using namespace mpdata;

namespace {

struct Timer {
    int calls = 0;
    double seconds = 0.0;                                  // summed, read intervals
    hipEvent_t open = nullptr;                             // pending begin
    std::vector<std::pair<hipEvent_t, hipEvent_t>> done;  // intervals not yet read
};

std::mutex g_mu;
bool g_on = false;
std::map<int, Timer> g_timers;
std::vector<hipEvent_t> g_pool;

hipEvent_t take_event() {
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return e;
}

// Adds the finished intervals of t to its sum and returns their events to
// the pool (caller holds g_mu).
hipError_t settle(Timer& t) {
    for (auto& p : t.done) {
        float ms = 0.f;
        hipError_t e = hipEventSynchronize(p.second);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, p.first, p.second);
        if (e != hipSuccess) return e;
        t.seconds += 1e-3 * (double)ms;
        g_pool.push_back(p.first);
        g_pool.push_back(p.second);
    }
    t.done.clear();
    return hipSuccess;
}

// The reference's labels for the ids this library records
// (common/mpi_parallel.f90:7088-7230).
const char* label(int id) {
    switch (id) {
        case 16: return "upwind_gpubc";
        case 51: return "adif_gau_gpubc";
        case 53: return "adif_std_gpubc";
        case 55: return "twos_gau_gpubc";
        case 57: return "twos_std_gpubc";
        case 300: return "eulagdwarf";
        case 301: return "timeloop";
        default: return "";
    }
}

}  // namespace

namespace mpdata {

bool tt_on() {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_on;
}

void tt_beg(int id, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_on) return;
    Timer& t = g_timers[id];
    if (t.open) g_pool.push_back(t.open);  // an unmatched begin is superseded
    t.open = take_event();
    if (t.open && hipEventRecord(t.open, s) != hipSuccess) {
        (void)hipGetLastError();
        g_pool.push_back(t.open);
        t.open = nullptr;
    }
}

void tt_end(int id, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_on) return;
    auto it = g_timers.find(id);
    if (it == g_timers.end() || !it->second.open) return;
    Timer& t = it->second;
    hipEvent_t e = take_event();
    if (!e || hipEventRecord(e, s) != hipSuccess) {
        (void)hipGetLastError();
        if (e) g_pool.push_back(e);
        g_pool.push_back(t.open);
        t.open = nullptr;
        return;
    }
    t.done.emplace_back(t.open, e);
    t.open = nullptr;
    t.calls++;
}

}  // namespace mpdata

extern "C" {

int mpdata_ttini(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto& kv : g_timers) {
        Timer& t = kv.second;
        for (auto& p : t.done) {
            (void)hipEventSynchronize(p.second);
            g_pool.push_back(p.first);
            g_pool.push_back(p.second);
        }
        if (t.open) g_pool.push_back(t.open);
    }
    g_timers.clear();
    g_on = true;
    return MPDATA_OK;
}

int mpdata_ttbeg(int id, void* stream) {
    if (id < 1 || id > 2000) return fail(MPDATA_EINVAL, "timer id %d (1..2000)", id);
    tt_beg(id, (hipStream_t)stream);
    return MPDATA_OK;
}

int mpdata_ttend(int id, void* stream) {
    if (id < 1 || id > 2000) return fail(MPDATA_EINVAL, "timer id %d (1..2000)", id);
    tt_end(id, (hipStream_t)stream);
    return MPDATA_OK;
}

int mpdata_ttget(int id, int* calls, double* seconds) {
    if (!calls || !seconds) return fail(MPDATA_EINVAL, "null output pointer");
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_timers.find(id);
    if (it == g_timers.end()) {
        *calls = 0;
        *seconds = 0.0;
        return MPDATA_OK;
    }
    HIPCHK(settle(it->second));
    *calls = it->second.calls;
    *seconds = it->second.seconds;
    return MPDATA_OK;
}

int mpdata_ttprt(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    std::vector<std::pair<double, int>> order;
    printf(" Pe #, Item #, Subrout,  # calls, .av timer,  mx timer,  mn timer, runit\n");
    for (auto& kv : g_timers) {
        Timer& t = kv.second;
        HIPCHK(settle(t));
        if (t.calls < 1) continue;
        // one process: the average, maximum and minimum over processes coincide
        printf("       %4d %4d    %-16s  %8d %12.4f %12.4f %12.4f\n", 0, kv.first,
               label(kv.first), t.calls, t.seconds, t.seconds, t.seconds);
        order.emplace_back(t.seconds, kv.first);
    }
    printf(" -------------------------------------------------\n");
    printf(" -----Sorted list of computing cost:--------------\n");
    printf(" -------------------------------------------------\n");
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<double, int>& a, const std::pair<double, int>& b) {
                         return a.first > b.first;
                     });
    for (auto& o : order) {
        const Timer& t = g_timers[o.second];
        printf("       %4d %4d    %-16s  %8d %12.4f %12.4f %12.4f\n", 0, o.second,
               label(o.second), t.calls, t.seconds, t.seconds, t.seconds);
    }
    fflush(stdout);
    return MPDATA_OK;
}

}  // extern "C"
