// Timers keyed by the reference's ids (mpdata_timers.hip): no-ops until
// mpdata_ttini turns them on.
#pragma once
#include <hip/hip_runtime.h>

namespace mpdata {
bool tt_on();
void tt_beg(int id, hipStream_t s);
void tt_end(int id, hipStream_t s);
}  // namespace mpdata
