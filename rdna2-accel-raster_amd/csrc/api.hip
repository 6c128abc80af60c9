// api.hip -- library-level entry points and error plumbing of librtbvh.so.
#include <cstdio>

#include "rt_internal.h"

namespace rt {

namespace {
thread_local char g_err[512] = "";
}

void set_error(const char* what, hipError_t e) {
    std::snprintf(g_err, sizeof g_err, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

// Launch errors (bad configuration, no device) surface here; asynchronous
// faults surface at the caller's next synchronisation.
rt_status launch_status(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(what, e);
        return RT_ERR_HIP;
    }
    return RT_OK;
}

}  // namespace rt

extern "C" const char* rt_version(void) { return "rtbvh 0.1 gfx950"; }

extern "C" const char* rt_last_error(void) { return rt::g_err; }

extern "C" rt_status rt_device_sync(void) {
    RT_TRY_HIP(hipDeviceSynchronize());
    return RT_OK;
}
