// Status reporting shared by the C ABI sources: every entry point returns an
// MPDATA_* code and leaves a message for mpdata_last_error() (per thread).
#pragma once
#include <hip/hip_runtime.h>

namespace mpdata {

// Records the formatted message and returns `code`.
int fail(int code, const char* fmt, ...);
// fail(MPDATA_EHIP, "<where>: <HIP error string>")
int hip_fail(hipError_t e, const char* where);
// The calling thread's last message.
const char* last_error();

}  // namespace mpdata

#define HIPCHK(call)                                               \
    do {                                                           \
        hipError_t e_ = (call);                                    \
        if (e_ != hipSuccess) return ::mpdata::hip_fail(e_, #call); \
    } while (0)
