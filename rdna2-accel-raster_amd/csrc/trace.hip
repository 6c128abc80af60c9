// trace.hip -- Box4 BVH traversal on gfx950: closest-hit and any-hit
// (occlusion) over rt_node4 / rt_tri64 trees, SoA rays.
//
// Follows benchmarks/exp0_probe/shaders/software_bvh_bvh4.comp:96-190: the
// same inverse-direction guard, slab test, child filter, child order (the
// shader's exchange sort by descending entry t, far children pushed first),
// leaf-first processing and Moller-Trumbore predicate, so on the same tree
// and rays the hit ID, t and box/triangle counters equal the CPU path bit
// for bit.  Unlike the shader the per-lane stack never drops entries: its
// depth is sized from the tree (3 * depth + 1 entries bound the live stack).
//
// Stack: LDS, entry-major ([S][block]) so lanes at the same depth hit
// distinct banks; trees deeper than the LDS variants use a private array.
#include "rt_internal.h"

namespace rt {
namespace {

constexpr int kBlock = 256;

struct Child {
    float key;
    int32_t child;
    int32_t count;
};

__device__ __forceinline__ void cx(Child& a, Child& b) {
    // software_bvh_bvh4.comp:151: swap when st[i] < st[j]
    if (a.key < b.key) {
        Child t = a;
        a = b;
        b = t;
    }
}

template <int S, bool ANY, bool COUNT, bool PRIVATE>
__global__ __launch_bounds__(kBlock) void k_trace(const rt_node4* __restrict__ nodes,
                                                  const rt_tri64* __restrict__ tris,
                                                  uint32_t node_count, rt_rays_soa rays,
                                                  uint32_t n, uint32_t* __restrict__ prim_out,
                                                  float* __restrict__ t_out,
                                                  uint32_t* __restrict__ bits_out,
                                                  uint32_t* __restrict__ box_out,
                                                  uint32_t* __restrict__ tri_out) {
    __shared__ int32_t lds_stack[PRIVATE ? 1 : S * kBlock];
    int32_t priv_stack[PRIVATE ? S : 1];
    const uint32_t tid = threadIdx.x;
    const uint32_t r = blockIdx.x * kBlock + tid;
    const bool valid = r < n;

    float ox = 0.0f, oy = 0.0f, oz = 0.0f, tmin = 1.0f;
    float dx = 0.0f, dy = 0.0f, dz = 1.0f, tmax = 0.0f;
    if (valid) {
        ox = rays.ox[r]; oy = rays.oy[r]; oz = rays.oz[r]; tmin = rays.tmin[r];
        dx = rays.dx[r]; dy = rays.dy[r]; dz = rays.dz[r]; tmax = rays.tmax[r];
    }
    // software_bvh_bvh4.comp:96-100
    const float ix = fabsf(dx) > 1e-6f ? 1.0f / dx : 1e6f;
    const float iy = fabsf(dy) > 1e-6f ? 1.0f / dy : 1e6f;
    const float iz = fabsf(dz) > 1e-6f ? 1.0f / dz : 1e6f;

    float hitT = tmax;
    uint32_t hitPrim = RT_MISS;
    uint32_t boxTests = 0, triTests = 0;
    bool done = false;

    auto st_at = [&](int i) -> int32_t& {
        if constexpr (PRIVATE) return priv_stack[i];
        else return lds_stack[i * kBlock + tid];
    };

    int ptr = 0;
    if (valid && node_count > 0) {
        st_at(0) = 0;
        ptr = 1;
    }

    while (ptr > 0) {
        --ptr;
        const int32_t ni = st_at(ptr);
        const float4* np = reinterpret_cast<const float4*>(nodes + ni);
        const float4 mnx = np[0], mxx = np[1], mny = np[2], mxy = np[3], mnz = np[4], mxz = np[5];
        const int4 ch = reinterpret_cast<const int4*>(np)[6];
        const int4 cn = reinterpret_cast<const int4*>(np)[7];
        if (COUNT) boxTests += 4;

        const float bminx[4] = {mnx.x, mnx.y, mnx.z, mnx.w};
        const float bmaxx[4] = {mxx.x, mxx.y, mxx.z, mxx.w};
        const float bminy[4] = {mny.x, mny.y, mny.z, mny.w};
        const float bmaxy[4] = {mxy.x, mxy.y, mxy.z, mxy.w};
        const float bminz[4] = {mnz.x, mnz.y, mnz.z, mnz.w};
        const float bmaxz[4] = {mxz.x, mxz.y, mxz.z, mxz.w};
        const int32_t chv[4] = {ch.x, ch.y, ch.z, ch.w};
        const int32_t cnv[4] = {cn.x, cn.y, cn.z, cn.w};

        Child e[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) e[p] = Child{-INFINITY, 0, 0};
        int c = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {  // :116-144 slab test + filter, kept in slot order
            float t0x = (bminx[k] - ox) * ix, t1x = (bmaxx[k] - ox) * ix;
            float t0y = (bminy[k] - oy) * iy, t1y = (bmaxy[k] - oy) * iy;
            float t0z = (bminz[k] - oz) * iz, t1z = (bmaxz[k] - oz) * iz;
            float tminX = fminf(t0x, t1x), tmaxX = fmaxf(t0x, t1x);
            float tminY = fminf(t0y, t1y), tmaxY = fmaxf(t0y, t1y);
            float tminZ = fminf(t0z, t1z), tmaxZ = fmaxf(t0z, t1z);
            float te = fmaxf(fmaxf(tminX, tminY), fmaxf(tminZ, tmin));
            float tx = fminf(fminf(tmaxX, tmaxY), fminf(tmaxZ, hitT));
            bool h = cnv[k] != 0 && tx >= te;
#pragma unroll
            for (int p = 0; p < 4; ++p)
                if (h && p == c) e[p] = Child{te, chv[k], cnv[k]};
            c += h ? 1 : 0;
        }
        if (c == 0) continue;
        // :149-156 exchange sort; padding (-inf) never moves
        cx(e[0], e[1]); cx(e[0], e[2]); cx(e[0], e[3]);
        cx(e[1], e[2]); cx(e[1], e[3]);
        cx(e[2], e[3]);

#pragma unroll
        for (int p = 0; p < 4; ++p) {  // :158-189
            if (done) break;
            if (e[p].count > 0) {
                const uint32_t first = (uint32_t)(~e[p].child);
                const uint32_t cnt = (uint32_t)e[p].count;
                if (COUNT) triTests += cnt;
                for (uint32_t q = 0; q < cnt; ++q) {
                    const float4* tp = reinterpret_cast<const float4*>(tris + first + q);
                    const float4 a = tp[0], b = tp[1], cc = tp[2];
                    float t;
                    if (!mt_hit(a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, cc.x, ox, oy, oz, dx, dy,
                                dz, &t))
                        continue;
                    if (t >= tmin && t < hitT) {
                        hitT = t;
                        hitPrim = tris[first + q].triangle_id;
                        if (ANY) {
                            done = true;
                            break;
                        }
                    }
                }
            } else if (e[p].count < 0) {
                st_at(ptr) = e[p].child;
                ++ptr;
            }
        }
        if (ANY && done) break;
    }

    if (ANY) {
        const uint64_t m = __ballot(valid && hitPrim != RT_MISS);
        const uint32_t lane = tid & 63u;
        const uint32_t r0 = r - lane;
        if (lane == 0 && r0 < n) bits_out[r0 >> 5] = (uint32_t)m;
        if (lane == 32 && r0 + 32 < n) bits_out[(r0 >> 5) + 1] = (uint32_t)(m >> 32);
        if (valid && prim_out) prim_out[r] = hitPrim;
    } else if (valid) {
        prim_out[r] = hitPrim;
        t_out[r] = hitT;
    }
    if (COUNT && valid) {
        box_out[r] = boxTests;
        tri_out[r] = triTests;
    }
}

template <int S, bool PRIVATE>
rt_status launch_s(hipStream_t s, const rt_bvh_dev* b, const rt_rays_soa* rays, uint32_t n,
                   uint32_t flags, rt_hits_soa* hits, uint32_t* bits, uint32_t* box,
                   uint32_t* tri) {
    const bool any = flags & RT_ANY;
    const bool cnt = flags & RT_COUNTERS;
    dim3 grid(ceil_div(n, kBlock)), block(kBlock);
    uint32_t* prim = hits ? hits->prim : nullptr;
    float* t = hits ? hits->t : nullptr;
#define RT_L(A, C)                                                                            \
    hipLaunchKernelGGL((k_trace<S, A, C, PRIVATE>), grid, block, 0, s, b->nodes, b->tris,     \
                       b->node_count, *rays, n, prim, t, bits, box, tri)
    if (any && cnt) RT_L(true, true);
    else if (any) RT_L(true, false);
    else if (cnt) RT_L(false, true);
    else RT_L(false, false);
#undef RT_L
    return launch_status("k_trace");
}

}  // namespace
}  // namespace rt

using namespace rt;

extern "C" rt_status rt_trace(rt_stream stream, const rt_bvh_dev* bvh, const rt_rays_soa* rays,
                              uint32_t n, uint32_t flags, rt_hits_soa* hits, uint32_t* occl_bits,
                              uint32_t* box_tests, uint32_t* tri_tests) {
    if (!bvh || !rays) return RT_ERR_ARG;
    if (flags & ~(RT_ANY | RT_COUNTERS)) return RT_ERR_ARG;
    if (n == 0) return RT_OK;
    if (!rays->ox || !rays->oy || !rays->oz || !rays->tmin || !rays->dx || !rays->dy ||
        !rays->dz || !rays->tmax)
        return RT_ERR_ARG;
    const bool any = flags & RT_ANY;
    if (any && !occl_bits) return RT_ERR_ARG;
    if (!any && (!hits || !hits->prim || !hits->t)) return RT_ERR_ARG;
    if ((flags & RT_COUNTERS) && (!box_tests || !tri_tests)) return RT_ERR_ARG;
    if (bvh->node_count > 0 && (!bvh->nodes || !bvh->tris)) return RT_ERR_ARG;
    const uint64_t need = 3ull * bvh->depth + 1;
    hipStream_t s = as_stream(stream);
    if (need <= 16) return launch_s<16, false>(s, bvh, rays, n, flags, hits, occl_bits, box_tests, tri_tests);
    if (need <= 24) return launch_s<24, false>(s, bvh, rays, n, flags, hits, occl_bits, box_tests, tri_tests);
    if (need <= 32) return launch_s<32, false>(s, bvh, rays, n, flags, hits, occl_bits, box_tests, tri_tests);
    if (need <= 48) return launch_s<48, false>(s, bvh, rays, n, flags, hits, occl_bits, box_tests, tri_tests);
    if (need <= 64) return launch_s<64, false>(s, bvh, rays, n, flags, hits, occl_bits, box_tests, tri_tests);
    if (need <= 256) return launch_s<256, true>(s, bvh, rays, n, flags, hits, occl_bits, box_tests, tri_tests);
    return RT_ERR_CAPACITY;
}
