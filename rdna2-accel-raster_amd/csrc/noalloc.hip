// noalloc.hip -- rt_gpu_build_noalloc: the device build of rt_gpu_build for
// callers that cannot allocate (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).  The build's scratch (sort
// buffers, radix-tree arrays, level boxes) lives in a file-scope device array
// of fixed size, one copy per GPU, held until the process ends; every call on
// a GPU shares that GPU's copy.  The build initialises every scratch word it
// reads before reading it, so nothing is cleared between calls.  Calls that
// overlap on one GPU (on different streams, or from several host threads)
// race on the array: serialise them.
#include <mutex>

#include "rt_internal.h"

namespace {

// Room for the largest build this entry point accepts: a RT_BUILD_RADIX tree
// of RT_NOALLOC_BUILD_MAX triangles needs 8,720,896 bytes, 10,031,616 with
// treelet restructuring (the implicit layout 1,642,752); the size is checked
// against the build's own plan on every call.
constexpr size_t kScratchBytes = 10u << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// The current device's address of `scratch` (hipGetSymbolAddress once per
// device, with that device current).  In host code the array's name is a
// host-side stand-in, not a device address: only this pointer is passed on.
rt_status device_scratch(void** out) {
    int dev = 0;
    RT_TRY_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return RT_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        RT_TRY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return RT_ERR_HIP;
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return RT_OK;
}

}  // namespace

extern "C" rt_status rt_gpu_build_noalloc(rt_stream stream, const float* d_pos,
                                          const uint32_t* d_idx, uint32_t n,
                                          const rt_build_cfg* cfg, rt_node4* d_nodes,
                                          uint32_t node_capacity, rt_tri64* d_tris,
                                          rt_build_info* d_info, uint32_t* d_sorted_ids,
                                          uint32_t* d_codes) {
    if (n > RT_NOALLOC_BUILD_MAX) return RT_ERR_CAPACITY;
    size_t need = 0;
    const rt_status q = rt_build_requirements(n, cfg, nullptr, &need);
    if (q != RT_OK) return q;
    if (need > kScratchBytes) return RT_ERR_CAPACITY;
    void* p = nullptr;
    if (need) {
        const rt_status st = device_scratch(&p);
        if (st != RT_OK) return st;
    }
    return rt_gpu_build(stream, d_pos, d_idx, n, cfg, d_nodes, node_capacity, d_tris, d_info,
                        d_sorted_ids, d_codes, p, need ? kScratchBytes : 0);
}
