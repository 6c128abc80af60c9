// rt_internal.h -- device helpers and host error plumbing shared by the
// librtbvh.so translation units.  gfx950 only; every file that includes this
// is compiled with -ffp-contract=off so float expressions round once per
// operator in the order written (the CPU path rounds the same way).
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>

#include <cfloat>
#include <cstdint>
#include <string>

#include "rt_bvh.h"

#pragma clang fp contract(off)

namespace rt {

// ---------------------------------------------------------------- host -----
void set_error(const char* what, hipError_t e);
rt_status launch_status(const char* what);

#define RT_TRY_HIP(call)                                   \
    do {                                                   \
        hipError_t _e = (call);                            \
        if (_e != hipSuccess) {                            \
            ::rt::set_error(#call, _e);                    \
            return RT_ERR_HIP;                             \
        }                                                  \
    } while (0)

inline hipStream_t as_stream(rt_stream s) { return reinterpret_cast<hipStream_t>(s); }

inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Deepest tree the traversal and refit entry points accept (node levels on a
// root-to-leaf path).  A radix tree over 30-bit codes plus 32-bit position
// tie-breaks is at most 62 binary levels deep, so its Box4 collapse fits;
// treelet restructuring can deepen a tree, and the build records the true
// depth in rt_build_info.depth (a deeper tree is rejected by the entry points
// given its depth).  A trace given depth 0 (unknown) sizes its stack for this
// bound and drops pushes beyond it instead of writing past the stack.
constexpr uint32_t kMaxDepth = 64;

// include/radv_bvh.hpp:193-215 compute_tree_layout (no 8-level cap).
struct TreeLayout {
    uint32_t levels = 0;
    uint32_t total_boxes = 0;
    uint64_t total_as_size = 0;
    uint32_t counts[32] = {0};
    uint32_t offsets[32] = {0};
};
TreeLayout tree_layout(uint32_t n);

// --------------------------------------------------------------- device ----

// software_bvh_bvh4.comp:50-54 pcg_hash.
__host__ __device__ inline uint32_t pcg_hash(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

// Uniform in [0,1) with 24 random bits (exact conversion and scaling).
__host__ __device__ inline float hash01(uint32_t seed, uint32_t i) {
    uint32_t h = pcg_hash(i * 0x9E3779B9u + seed);
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}

// Ray i of rt_gen_random_rays(seed, lo, hi, tmax): origin uniform in the box
// [lo, hi), direction uniform on the unit sphere, tmin 0.  Shared by the
// one-request generator and the batched one, so both give the same rays.
__device__ __forceinline__ void random_ray(uint32_t seed, uint32_t i, const float lo[3],
                                           const float hi[3], float o[3], float d[3]) {
    const float u0 = hash01(seed, i * 5 + 0), u1 = hash01(seed, i * 5 + 1);
    const float u2 = hash01(seed, i * 5 + 2), u3 = hash01(seed, i * 5 + 3);
    const float u4 = hash01(seed, i * 5 + 4);
    const float z = 1.0f - 2.0f * u3;
    const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    const float phi = 6.2831853f * u4;
    o[0] = lo[0] + u0 * (hi[0] - lo[0]);
    o[1] = lo[1] + u1 * (hi[1] - lo[1]);
    o[2] = lo[2] + u2 * (hi[2] - lo[2]);
    d[0] = r * cosf(phi);
    d[1] = r * sinf(phi);
    d[2] = z;
}

// Quad q (two triangles, 18 floats at o) of include/radv_bvh.hpp:224-283
// generate_dynamic_grid_mesh over a cols x rows grid of [-extent, extent]^2,
// z = 5 + 0.2 sin(2 x + t) cos(2 y + t).
__device__ __forceinline__ void dyn_grid_quad(uint32_t cols, uint32_t rows, float t,
                                              float extent, uint32_t q, float* o) {
    const uint32_t r = q / cols, c = q % cols;
    const float dx = (2.0f * extent) / (float)cols, dy = (2.0f * extent) / (float)rows;
    const float x0 = -extent + (float)c * dx, x1 = x0 + dx;
    const float y0 = -extent + (float)r * dy, y1 = y0 + dy;
    auto z = [t](float x, float y) {
        return 5.0f + 0.2f * sinf(x * 2.0f + t) * cosf(y * 2.0f + t);
    };
    const float z00 = z(x0, y0), z10 = z(x1, y0), z01 = z(x0, y1), z11 = z(x1, y1);
    const float tv[18] = {x0, y0, z00, x1, y0, z10, x0, y1, z01,
                          x1, y0, z10, x1, y1, z11, x0, y1, z01};
#pragma unroll
    for (int k = 0; k < 18; ++k) o[k] = tv[k];
}

// include/radv_bvh.hpp:228-233 grid shape of n triangles: quads = n / 2,
// cols the largest divisor of quads not above sqrt(quads).
inline void dyn_grid_shape(uint32_t n, uint32_t* cols, uint32_t* rows) {
    const uint32_t quads = n / 2;
    uint32_t c = (uint32_t)std::sqrt((double)quads);
    while (c > 1 && quads % c != 0) c--;
    *cols = c;
    *rows = c ? quads / c : 0;
}

// Batched payload generators (gen.hip) for the submit entry points.
struct GridGen {  // one deforming grid: 9 floats per triangle into xyz
    float* xyz;
    uint32_t cols, rows;
    float t, extent;
};
rt_status launch_grid_gen(hipStream_t s, const GridGen* d_gen, uint32_t count,
                          uint32_t max_quads);
rt_status launch_ray_gen(hipStream_t s, const rt_trace_item* d_items, const rt_ray_gen* d_gen,
                         const uint32_t* d_chunk_start, uint32_t count, uint32_t total_chunks,
                         uint32_t chunk_rays);

// Moller-Trumbore with the predicates of software_bvh_bvh4.comp:166-182:
// |det| < 1e-7 rejects, u,v inclusive in [0,1], u+v <= 1.
__device__ __forceinline__ bool mt_hit(float v0x, float v0y, float v0z, float v1x, float v1y,
                                       float v1z, float v2x, float v2y, float v2z, float ox,
                                       float oy, float oz, float dx, float dy, float dz,
                                       float* tout, float eps = 1e-7f) {
    float e1x = v1x - v0x, e1y = v1y - v0y, e1z = v1z - v0z;
    float e2x = v2x - v0x, e2y = v2y - v0y, e2z = v2z - v0z;
    float hx = dy * e2z - dz * e2y;
    float hy = dz * e2x - dx * e2z;
    float hz = dx * e2y - dy * e2x;
    float a = e1x * hx + e1y * hy + e1z * hz;
    if (fabsf(a) < eps) return false;
    float f = 1.0f / a;
    float sx = ox - v0x, sy = oy - v0y, sz = oz - v0z;
    float u = f * (sx * hx + sy * hy + sz * hz);
    if (u < 0.0f || u > 1.0f) return false;
    float qx = sy * e1z - sz * e1y;
    float qy = sz * e1x - sx * e1z;
    float qz = sx * e1y - sy * e1x;
    float v = f * (dx * qx + dy * qy + dz * qz);
    if (v < 0.0f || u + v > 1.0f) return false;
    *tout = f * (e2x * qx + e2y * qy + e2z * qz);
    return true;
}

// Workgroup barrier that orders LDS only.  __syncthreads() also waits for
// every outstanding global store (gfx9 counts loads and stores on one
// counter), which a kernel that never reads its own global writes back does
// not need.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Orders this wave's LDS accesses before and after (a wave runs its LDS
// instructions in order; this stops the compiler from moving them).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// x / 3 rounded to nearest without the division sequence: q = x * RN(1/3),
// one residual step with the exact FMA residual x - 3q, sign taken from x
// (the step turns -0 into +0).  Bit-identical to IEEE x / 3.0f for every
// finite float (all 2^32 patterns checked on the host; tests repeat a
// sample of that check).
__host__ __device__ inline float div3(float x) {
    const float y = 0x1.555556p-2f;  // RN(1/3)
    const float q = x * y;
    const float r = fmaf(-3.0f, q, x);
    return copysignf(fmaf(r, y, q), x);
}

// Float <-> order-preserving uint (for atomicMin / atomicMax on bounds).
__device__ __forceinline__ uint32_t f2o(float f) {
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t o) {
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

// Spread the low 10 bits of v to every third bit.
__host__ __device__ inline uint32_t expand_bits10(uint32_t v) {
    v &= 0x3FFu;
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

}  // namespace rt
