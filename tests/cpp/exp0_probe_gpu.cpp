// exp0_probe_gpu -- the reference's exp0 traversal experiment
// (benchmarks/exp0_probe/probe_main.cpp, results/exp0_ray_stats.csv:2-7)
// driven from C++ through the rtgpu host interface (gpu_bvh.hpp) on the GPU:
// the reference terrain (common/mesh_gen.hpp createTessellatedTerrain, as
// its -O3 -march=native build evaluates it), the reference's SoftwareBVH4
// built on the host (the oracle's byte-exact restatement, test
// infrastructure) and uploaded with GpuBVH4::fromHost, the reference's
// procedural coherent / incoherent rays (1 << 20, tmin 0.01, tmax 1000),
// generated on the device (rt_gen_exp0_rays; checked bit for bit against the
// oracle's restatement on the first terrain).
// Prints the CSV's software-path columns plus this GPU's traversal rate:
//   Triangles,RayPattern,HitRate_SW_pct,AvgBoxTests_PerRay,AvgBoxTests_HitOnly,
//   AvgTriTests_PerRay,AvgTriTests_HitOnly,GPU_MRays_per_sec
// BVH2 rows (prefix "bvh2,"): the same terrain's SoftwareBVH (before the
// collapse) traced with the loop of software_bvh_procedural.comp
// (GpuBVH2, tmin 0.01) and of its LDS-stack variant software_bvh_lds.comp
// (tmin 0.001): Triangles,RayPattern,Shader,HitRate_pct,AvgBoxTests,
// AvgTriTests,GPU_MRays_per_sec.
// Then builds the same terrain on the device (GpuBVH4::build, radix),
// refits it at the same positions and checks the bytes are unchanged; for
// the smallest terrain also the RADV image of its Morton order, traced over
// its bytes (RadvBVH), against the implicit Box4 build of the same mesh, and
// the image metrics of the two hit-distance images (identical: mse 0).
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

std::string fmt(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%.6f", v);
    return b;
}

// The device-generated exp0 rays equal the oracle's restatement of the shader's
// generator (software_bvh_bvh4.comp:50-94) bit for bit.  Returns false on a mismatch.
bool rays_match_oracle(const rtgpu::Rays& rays, uint32_t n, int coherent, const float cam[3],
                       const float look[3], hipStream_t s) {
    std::vector<float> f[6];
    for (auto& v : f) v.assign(n, 0.0f);
    or_exp0_rays(n, coherent, cam, look, f[0].data(), f[1].data(), f[2].data(), f[3].data(),
                 f[4].data(), f[5].data());
    const int field[6] = {0, 1, 2, 4, 5, 6};
    for (int k = 0; k < 6; ++k) {
        const std::vector<float> d = rays.download(field[k], s);
        if (std::memcmp(d.data(), f[k].data(), (size_t)n * 4) != 0) {
            std::printf("# device exp0 rays (%s) differ from the oracle in field %d\n",
                        coherent ? "coherent" : "incoherent", field[k]);
            return false;
        }
    }
    std::printf("# device exp0 rays (%s) equal the oracle's bit for bit\n",
                coherent ? "coherent" : "incoherent");
    return true;
}

// The RADV image of the mesh (Morton order) traced over its bytes equals the implicit
// Box4 build of the same mesh; the image metrics of the two hit-distance images are
// those of identical images.  Returns 1 on a mismatch.
int radv_check(const SceneMesh& mesh, hipStream_t s) {
    rtgpu::RadvBVH img = rtgpu::RadvBVH::build(mesh, RT_RADV_MORTON, s);
    rtgpu::GpuBVH4 box = rtgpu::GpuBVH4::build(mesh, 4, RT_BUILD_IMPLICIT, s);
    const rt_radv_info info = img.describe();
    const uint32_t w = 256, h = 256, n = w * h;
    const float cam[3] = {0.0f, 50.0f, -45.0f}, look[3] = {0.0f, 0.0f, 5.0f};
    rtgpu::Rays rays = rtgpu::Rays::exp0(n, false, cam, look, 0.01f, 1000.0f, s);
    rtgpu::Hits a, b;
    img.intersect(rays, a, true);
    box.intersect(rays, b, true);
    const bool same = a.prim.download(s) == b.prim.download(s) &&
                      a.boxTests.download(s) == b.boxTests.download(s) &&
                      a.triTests.download(s) == b.triTests.download(s);
    const rtgpu::ImageQuality q = rtgpu::imageMetrics(a.t.data(), b.t.data(), w, h, s);
    std::printf("# RADV image: %u box nodes, %llu triangles, bvh_offset 0x%x; traced over its bytes "
                "%s the implicit Box4 tree; metrics of the t images: mse %g psnr %g ssim %g\n",
                info.box_count, (unsigned long long)info.primitive_count, info.bvh_offset,
                same ? "equals" : "DIFFERS FROM", q.mse, q.psnr, q.ssim);
    return same && q.mse == 0.0 && q.psnr == 100.0 ? 0 : 1;
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
    std::vector<std::string> bvh2_rows;
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
        rtgpu::GpuBVH2 bvh2 = rtgpu::GpuBVH2::fromHost(n2.data(), c2, t48.data(), order.data(),
                                                       ntri, s);
        for (int coherent = 1; coherent >= 0; --coherent) {
            const uint32_t n = 1u << 20;
            const float cam[3] = {0.0f, 50.0f, -45.0f}, look[3] = {0.0f, 0.0f, 5.0f};
            // the probe's procedural rays, generated on the device (rt_gen_exp0_rays)
            rtgpu::Rays rays = rtgpu::Rays::exp0(n, coherent != 0, cam, look, 0.01f, 1000.0f, s);
            if (target == targets.front() && !rays_match_oracle(rays, n, coherent, cam, look, s))
                return 1;
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
            // occluder hints: each ray's own closest hit (period 0) and the hits of the
            // first quarter of the rays for every quarter (period n / 4); the bits equal
            // the plain any-hit traversal's either way
            {
                rtgpu::DeviceArray<uint32_t> plain, hinted;
                bvh.occluded(rays, plain);
                const std::vector<uint32_t> want = plain.download(s);
                for (uint32_t period : {0u, n / 4}) {
                    for (int wave = 0; wave < 2; ++wave) {
                        bvh.occludedHinted(rays, hits.prim, period, hinted, wave != 0);
                        if (hinted.download(s) != want) {
                            std::printf("# hinted occlusion differs: %u triangles, period %u, "
                                        "wave %d\n", ntri, period, wave);
                            return 1;
                        }
                    }
                }
                std::printf("# hinted occlusion equals any hit: %u triangles, %s\n", ntri,
                            coherent ? "Coherent" : "Incoherent");
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
            // the BVH2 shaders on the same rays: tmin 0.01 (scratch stack) and 0.001 (LDS)
            for (int lds = 0; lds < 2; ++lds) {
                rtgpu::Rays r2 = rtgpu::Rays::exp0(n, coherent != 0, cam, look,
                                                   lds ? 0.001f : 0.01f, 1000.0f, s);
                rtgpu::Hits h2;
                bvh2.intersect(r2, h2, true);
                const std::vector<uint32_t> p2 = h2.prim.download(s);
                const std::vector<uint32_t> b2 = h2.boxTests.download(s);
                const std::vector<uint32_t> q2 = h2.triTests.download(s);
                double hits2 = 0, box2 = 0, tri2 = 0;
                for (uint32_t r = 0; r < n; ++r) {
                    hits2 += p2[r] != RT_MISS;
                    box2 += b2[r];
                    tri2 += q2[r];
                }
                bvh2.intersect(r2, h2);
                rtgpu::hip_check(hipEventCreate(&e0), "hipEventCreate");
                rtgpu::hip_check(hipEventCreate(&e1), "hipEventCreate");
                rtgpu::hip_check(hipEventRecord(e0, s), "hipEventRecord");
                for (int k = 0; k < 10; ++k) bvh2.intersect(r2, h2);
                rtgpu::hip_check(hipEventRecord(e1, s), "hipEventRecord");
                rtgpu::hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
                float ms2 = 0.0f;
                rtgpu::hip_check(hipEventElapsedTime(&ms2, e0, e1), "hipEventElapsedTime");
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                bvh2_rows.push_back(std::to_string(ntri) + "," +
                                    (coherent ? "Coherent" : "Incoherent") + "," +
                                    (lds ? "Software_ALU_LDS" : "Software_ALU_Scratch") + "," +
                                    fmt(100.0 * hits2 / n) + "," + fmt(box2 / n) + "," +
                                    fmt(tri2 / n) + "," + fmt(10.0 * n / (ms2 * 1e-3) / 1e6));
            }
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
        if (target == targets.front() && ntri <= 4096 && radv_check(mesh, s)) return 1;
    }
    for (const std::string& row : bvh2_rows) std::printf("bvh2,%s\n", row.c_str());
    (void)hipStreamDestroy(s);
    return 0;
}
