// exp0_probe_gpu -- the reference's exp0 traversal experiment
// (benchmarks/exp0_probe/probe_main.cpp, results/exp0_ray_stats.csv:2-7)
// driven from C++ through the rtgpu host interface (gpu_bvh.hpp) on the GPU:
// the reference terrain (common/mesh_gen.hpp createTessellatedTerrain, as
// its -O3 -march=native build evaluates it), the reference's SoftwareBVH4
// built on the host (the oracle's byte-exact restatement, test
// infrastructure) and uploaded with GpuBVH4::fromHost, the reference's
// procedural coherent / incoherent rays (1 << 20, tmin 0.01, tmax 1000).
// Prints the CSV's software-path columns plus this GPU's traversal rate:
//   Triangles,RayPattern,HitRate_SW_pct,AvgBoxTests_PerRay,AvgBoxTests_HitOnly,
//   AvgTriTests_PerRay,AvgTriTests_HitOnly,GPU_MRays_per_sec
// Then builds the same terrain on the device (GpuBVH4::build, radix),
// refits it at the same positions and checks the bytes are unchanged.
//
//   exp0_probe_gpu [targets...]   (default 1000 100000 1000000)
//   exp0_probe_gpu --host-only    argument checks of the wrapper, no device work
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gpu_bvh.hpp"
#include "oracle.h"

namespace {

struct Vertex {
    float pos[3];
    float normal[3];
};
struct SceneMesh {  // the shape of common/mesh_gen.hpp:12-24
    std::vector<Vertex> vertices;
    std::vector<uint32_t> indices;
};

SceneMesh terrain(uint32_t target, std::vector<float>& pos, std::vector<uint32_t>& idx) {
    const uint32_t res = or_terrain_res(target);
    pos.assign((size_t)res * res * 3, 0.0f);
    idx.assign((size_t)(res - 1) * (res - 1) * 6, 0u);
    or_terrain_ex(target, 100.0f, 1, pos.data(), idx.data());
    SceneMesh m;
    m.vertices.resize((size_t)res * res);
    for (size_t v = 0; v < m.vertices.size(); ++v)
        for (int c = 0; c < 3; ++c) m.vertices[v].pos[c] = pos[3 * v + c];
    m.indices = idx;
    return m;
}

int host_only() {
    // rt_status errors surface as rtgpu::Error before any device work
    int fails = 0;
    try {
        rt_build_cfg cfg = {7u, 4u, 0u, 0u};
        rtgpu::check(rt_build_requirements(10, &cfg, nullptr, nullptr), "rt_build_requirements");
        std::printf("FAIL: bad algorithm accepted\n");
        ++fails;
    } catch (const rtgpu::Error& e) {
        std::printf("ok: %s\n", e.what());
    }
    uint32_t cap = 0;
    size_t scratch = 0;
    rt_build_cfg cfg = {RT_BUILD_RADIX, 4u, 0u, 0u};
    rtgpu::check(rt_build_requirements(2500000, &cfg, &cap, &scratch), "rt_build_requirements");
    if (cap != 2499999u) {
        std::printf("FAIL: capacity %u\n", cap);
        ++fails;
    }
    std::printf("host-only %s\n", fails ? "FAILED" : "ok");
    return fails ? 1 : 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc > 1 && std::strcmp(argv[1], "--host-only") == 0) return host_only();
    std::vector<uint32_t> targets;
    for (int i = 1; i < argc; ++i) targets.push_back((uint32_t)std::strtoul(argv[i], nullptr, 10));
    if (targets.empty()) targets = {1000u, 100000u, 1000000u};
    hipStream_t s;
    rtgpu::hip_check(hipStreamCreate(&s), "hipStreamCreate");
    std::printf("Triangles,RayPattern,HitRate_SW_pct,AvgBoxTests_PerRay,AvgBoxTests_HitOnly,"
                "AvgTriTests_PerRay,AvgTriTests_HitOnly,GPU_MRays_per_sec\n");
    for (uint32_t target : targets) {
        std::vector<float> pos;
        std::vector<uint32_t> idx;
        const SceneMesh mesh = terrain(target, pos, idx);
        const uint32_t ntri = (uint32_t)(idx.size() / 3);
        // the reference's SoftwareBVH::build + SoftwareBVH4::fromBVH2 on the host
        std::vector<or_bvh2_node> n2((size_t)2 * ntri + 1);
        std::vector<float> t48((size_t)ntri * 12);
        std::vector<uint32_t> order(ntri);
        uint32_t c2 = 0, c4 = 0;
        if (or_bvh2_build(pos.data(), (uint32_t)(pos.size() / 3), idx.data(), ntri, 4, n2.data(),
                          (uint32_t)n2.size(), &c2, t48.data(), order.data()))
            throw rtgpu::Error("or_bvh2_build");
        std::vector<or_bvh4_node> n4(c2);
        if (or_bvh4_collapse(n2.data(), c2, n4.data(), c2, &c4)) throw rtgpu::Error("collapse");
        const uint32_t depth = or_bvh4_depth(n4.data(), c4);
        rtgpu::GpuBVH4 bvh = rtgpu::GpuBVH4::fromHost(n4.data(), c4, t48.data(), order.data(),
                                                      ntri, depth, s);
        for (int coherent = 1; coherent >= 0; --coherent) {
            const uint32_t n = 1u << 20;
            std::vector<float> f[8];
            for (auto& v : f) v.assign(n, 0.0f);
            const float cam[3] = {0.0f, 50.0f, -45.0f}, look[3] = {0.0f, 0.0f, 5.0f};
            or_exp0_rays(n, coherent, cam, look, f[0].data(), f[1].data(), f[2].data(),
                         f[4].data(), f[5].data(), f[6].data());
            for (uint32_t r = 0; r < n; ++r) {
                f[3][r] = 0.01f;
                f[7][r] = 1000.0f;
            }
            const float* p[8];
            for (int k = 0; k < 8; ++k) p[k] = f[k].data();
            rtgpu::Rays rays(p, n, s);
            rtgpu::Hits hits;
            bvh.intersect(rays, hits, true);
            const std::vector<uint32_t> prim = hits.prim.download(s);
            const std::vector<uint32_t> box = hits.boxTests.download(s);
            const std::vector<uint32_t> tri = hits.triTests.download(s);
            double nh = 0, sb = 0, st = 0, sbh = 0, sth = 0;
            for (uint32_t r = 0; r < n; ++r) {
                sb += box[r];
                st += tri[r];
                if (prim[r] != RT_MISS) {
                    nh += 1;
                    sbh += box[r];
                    sth += tri[r];
                }
            }
            // traversal rate: the plain closest-hit launch, events around 10 launches
            hipEvent_t e0, e1;
            rtgpu::hip_check(hipEventCreate(&e0), "hipEventCreate");
            rtgpu::hip_check(hipEventCreate(&e1), "hipEventCreate");
            bvh.intersect(rays, hits);
            rtgpu::hip_check(hipEventRecord(e0, s), "hipEventRecord");
            for (int k = 0; k < 10; ++k) bvh.intersect(rays, hits);
            rtgpu::hip_check(hipEventRecord(e1, s), "hipEventRecord");
            rtgpu::hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
            float ms = 0.0f;
            rtgpu::hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            std::printf("%u,%s,%.6f,%.6f,%.6f,%.6f,%.6f,%.1f\n", ntri,
                        coherent ? "Coherent" : "Incoherent", 100.0 * nh / n, sb / n,
                        nh ? sbh / nh : 0.0, st / n, nh ? sth / nh : 0.0,
                        10.0 * n / (ms * 1e-3) / 1e6);
        }
        // device build of the same mesh, refit at the same positions: unchanged bytes
        rtgpu::GpuBVH4 dev = rtgpu::GpuBVH4::build(mesh, 4, RT_BUILD_RADIX, s);
        const std::vector<rt_node4> before = dev.downloadNodes();
        const std::vector<rt_tri64> tb = dev.downloadTriangles();
        dev.refit(mesh);
        const std::vector<rt_node4> after = dev.downloadNodes();
        const std::vector<rt_tri64> ta = dev.downloadTriangles();
        const bool same = std::memcmp(before.data(), after.data(),
                                      before.size() * sizeof(rt_node4)) == 0 &&
                          std::memcmp(tb.data(), ta.data(), tb.size() * sizeof(rt_tri64)) == 0;
        std::printf("# device radix build: %u nodes, depth %u; refit at the same positions %s\n",
                    dev.nodeCount(), dev.depth(), same ? "unchanged" : "CHANGED");
        if (!same) return 1;
    }
    (void)hipStreamDestroy(s);
    return 0;
}
