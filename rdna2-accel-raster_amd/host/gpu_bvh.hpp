// gpu_bvh.hpp -- C++ host interface over librtbvh.so (include/rt_bvh.h),
// keeping the build / intersect interface of the reference's
// common/software_bvh.hpp:
//
//   SoftwareBVH::build(mesh, maxPrimsPerLeaf) + SoftwareBVH4::fromBVH2
//       (software_bvh.hpp:35,193)             -> rtgpu::GpuBVH4::build(mesh, maxPrimsPerLeaf)
//   a host-built SoftwareBVH4 uploaded as SSBOs -> rtgpu::GpuBVH4::fromHost(...)
//   software_bvh_bvh4.comp closest hit          -> GpuBVH4::intersect
//   TerminateOnFirstHit ray queries (rq_ao.comp) -> GpuBVH4::occluded
//   the same with occluder hints (expB_shadows)   -> GpuBVH4::occludedHinted
//   (no reference counterpart)                  -> GpuBVH4::refit
//   a host-built SoftwareBVH + software_bvh_procedural.comp (BVH2 loop)
//                                               -> rtgpu::GpuBVH2::fromHost, intersect
//   fast_bvh_builder image + ray queries over it -> rtgpu::RadvBVH::build, intersect
//   dump_bvh_structure (radv_bvh.hpp:311-390)    -> RadvBVH::describe
//   ImageMetrics::compute (image_metrics.hpp:16) -> rtgpu::imageMetrics (device images)
//
// Header-only.  Device memory is owned by RAII arrays; every library status
// other than RT_OK, and every failing HIP call, throws rtgpu::Error (the
// reference's setup code throws std::runtime_error the same way).  All work
// is asynchronous on the object's HIP stream.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "rt_bvh.h"

namespace rtgpu {

struct Error : std::runtime_error {
    explicit Error(const std::string& what) : std::runtime_error(what) {}
};

inline void check(rt_status s, const char* what) {
    if (s == RT_OK) return;
    std::string m = std::string(what) + " failed: rt_status " + std::to_string((int)s);
    if (s == RT_ERR_HIP) m += std::string(" (") + rt_last_error() + ")";
    throw Error(m);
}

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

// Device array of n elements of T (uninitialised), freed with the object.
template <class T>
class DeviceArray {
public:
    DeviceArray() = default;
    explicit DeviceArray(size_t n) { resize(n); }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    DeviceArray(DeviceArray&& o) noexcept : p_(std::exchange(o.p_, nullptr)), n_(o.n_) {}
    DeviceArray& operator=(DeviceArray&& o) noexcept {
        if (this != &o) {
            release();
            p_ = std::exchange(o.p_, nullptr);
            n_ = o.n_;
        }
        return *this;
    }
    ~DeviceArray() { release(); }

    void resize(size_t n) {
        if (n == n_ && p_) return;
        release();
        n_ = n;
        if (n) hip_check(hipMalloc(&p_, n * sizeof(T)), "hipMalloc");
    }
    void upload(const T* src, size_t n, hipStream_t s) {
        resize(n);
        if (n) hip_check(hipMemcpyAsync(p_, src, n * sizeof(T), hipMemcpyHostToDevice, s),
                         "hipMemcpyAsync H2D");
    }
    std::vector<T> download(hipStream_t s, size_t n = SIZE_MAX) const {
        std::vector<T> h(n == SIZE_MAX ? n_ : n);
        if (!h.empty())
            hip_check(hipMemcpyAsync(h.data(), p_, h.size() * sizeof(T), hipMemcpyDeviceToHost, s),
                      "hipMemcpyAsync D2H");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        return h;
    }
    T* data() const { return p_; }
    size_t size() const { return n_; }

private:
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* p_ = nullptr;
    size_t n_ = 0;
};

// Positions of a SceneMesh-shaped mesh (common/mesh_gen.hpp:12-24: vertices[i].pos[3]
// and a flat index list), packed as float3 for the device.
template <class Mesh>
std::vector<float> packed_positions(const Mesh& mesh) {
    std::vector<float> pos(mesh.vertices.size() * 3);
    for (size_t v = 0; v < mesh.vertices.size(); ++v)
        for (int c = 0; c < 3; ++c) pos[3 * v + c] = mesh.vertices[v].pos[c];
    return pos;
}

// The reference probe's ray record (benchmarks/exp0_probe/shaders/software_bvh.comp:5-9):
// origin.w = tmin, direction.w = tmax.
struct RayInput {
    float origin[4];
    float direction[4];
};

// Structure-of-arrays rays in device memory (32 B per ray).
class Rays {
public:
    Rays() = default;
    Rays(const float* const host[8], size_t n, hipStream_t s) {
        for (int k = 0; k < 8; ++k) a_[k].upload(host[k], n, s);
        n_ = n;
    }
    static Rays fromRayInputs(const std::vector<RayInput>& in, hipStream_t s) {
        std::vector<float> f[8];
        for (auto& v : f) v.resize(in.size());
        for (size_t r = 0; r < in.size(); ++r) {
            f[0][r] = in[r].origin[0];
            f[1][r] = in[r].origin[1];
            f[2][r] = in[r].origin[2];
            f[3][r] = in[r].origin[3];
            f[4][r] = in[r].direction[0];
            f[5][r] = in[r].direction[1];
            f[6][r] = in[r].direction[2];
            f[7][r] = in[r].direction[3];
        }
        const float* p[8];
        for (int k = 0; k < 8; ++k) p[k] = f[k].data();
        return Rays(p, in.size(), s);
    }
    // The exp0 probe's procedural rays generated on the device
    // (software_bvh_bvh4.comp:50-94, rt_gen_exp0_rays): a 1024-wide camera grid
    // from cam towards look (coherent) or pcg_hash origins / directions.
    static Rays exp0(size_t n, bool coherent, const float cam[3], const float look[3],
                     float tmin, float tmax, hipStream_t s) {
        Rays r;
        for (auto& a : r.a_) a.resize(n);
        r.n_ = n;
        const rt_rays_out o = {r.a_[0].data(), r.a_[1].data(), r.a_[2].data(), r.a_[3].data(),
                               r.a_[4].data(), r.a_[5].data(), r.a_[6].data(), r.a_[7].data()};
        check(rt_gen_exp0_rays(s, (uint32_t)n, coherent ? 1 : 0, cam, look, tmin, tmax, &o),
              "rt_gen_exp0_rays");
        return r;
    }
    rt_rays_soa view() const {
        return {a_[0].data(), a_[1].data(), a_[2].data(), a_[3].data(),
                a_[4].data(), a_[5].data(), a_[6].data(), a_[7].data()};
    }
    std::vector<float> download(int field, hipStream_t s) const { return a_[field].download(s); }
    size_t size() const { return n_; }

private:
    DeviceArray<float> a_[8];
    size_t n_ = 0;
};

// Per-ray results of intersect (closest hit) in device memory.
struct Hits {
    DeviceArray<uint32_t> prim;  // original triangle index, RT_MISS on a miss
    DeviceArray<float> t;        // hit distance (tmax on a miss)
    DeviceArray<uint32_t> boxTests, triTests;  // with counters only
};

// Device-resident Box4 BVH (GpuBvh4Node nodes, 64-B leaves carrying triangle IDs).
class GpuBVH4 {
public:
    // SoftwareBVH::build(mesh, maxPrimsPerLeaf) + SoftwareBVH4::fromBVH2: a device build
    // of the same Box4 format (RT_BUILD_RADIX: Morton radix tree collapsed with
    // fromBVH2's rule; RT_BUILD_IMPLICIT: the fast_bvh_builder tree).  treeletRounds
    // (RT_BUILD_RADIX only): rounds of SAH treelet restructuring before the collapse.
    template <class Mesh>
    static GpuBVH4 build(const Mesh& mesh, uint32_t maxPrimsPerLeaf = 4,
                         uint32_t algorithm = RT_BUILD_RADIX, hipStream_t s = nullptr,
                         uint32_t treeletRounds = 0) {
        const std::vector<float> pos = packed_positions(mesh);
        GpuBVH4 b;
        b.s_ = s;
        b.pos_.upload(pos.data(), pos.size(), s);
        b.idx_.upload(mesh.indices.data(), mesh.indices.size(), s);
        b.build_device((uint32_t)(mesh.indices.size() / 3), maxPrimsPerLeaf, algorithm,
                       treeletRounds);
        return b;
    }

    // A host-built SoftwareBVH4 (software_bvh.hpp:189-291): its node array uploads
    // unchanged (GpuBvh4Node == rt_node4); its GpuTriangle array (48 B each, leaf order)
    // plus the leaf-order original triangle indices become 64-B leaves.
    static GpuBVH4 fromHost(const void* nodes, uint32_t nodeCount, const float* tris48,
                            const uint32_t* leafOrder, uint32_t triCount, uint32_t depth,
                            hipStream_t s = nullptr) {
        GpuBVH4 b;
        b.s_ = s;
        b.nodes_.resize(nodeCount);
        if (nodeCount)
            hip_check(hipMemcpyAsync(b.nodes_.data(), nodes, (size_t)nodeCount * sizeof(rt_node4),
                                     hipMemcpyHostToDevice, s),
                      "hipMemcpyAsync nodes");
        DeviceArray<float> t48;
        DeviceArray<uint32_t> order;
        t48.upload(tris48, (size_t)triCount * 12, s);
        if (leafOrder) order.upload(leafOrder, triCount, s);
        b.tris_.resize(triCount);
        check(rt_convert_tris48(s, t48.data(), leafOrder ? order.data() : nullptr, triCount,
                                b.tris_.data()),
              "rt_convert_tris48");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");  // t48 / order freed next
        b.nodeCount_ = nodeCount;
        b.triCount_ = triCount;
        b.depth_ = depth;
        return b;
    }

    // software_bvh_bvh4.comp: closest hit of every ray (bit-identical to the CPU path
    // on the same tree).  counters: also the per-ray box / triangle test counts.
    void intersect(const Rays& rays, Hits& out, bool counters = false,
                   const rt_trace_opts* opts = nullptr) const {
        const size_t n = rays.size();
        out.prim.resize(n);
        out.t.resize(n);
        if (counters) {
            out.boxTests.resize(n);
            out.triTests.resize(n);
        }
        const rt_bvh_dev v = view();
        const rt_rays_soa r = rays.view();
        rt_hits_soa h = {out.prim.data(), out.t.data()};
        check(rt_trace_ex(s_, &v, &r, (uint32_t)n, RT_CLOSEST | (counters ? RT_COUNTERS : 0u),
                          opts, &h, nullptr, counters ? out.boxTests.data() : nullptr,
                          counters ? out.triTests.data() : nullptr),
              "rt_trace_ex (closest)");
    }

    // Any hit (TerminateOnFirstHit): bit r of word r / 32 = ray r occluded.
    void occluded(const Rays& rays, DeviceArray<uint32_t>& bits,
                  const rt_trace_opts* opts = nullptr) const {
        const size_t n = rays.size();
        bits.resize(n ? (n + 31) / 32 : 1);
        const rt_bvh_dev v = view();
        const rt_rays_soa r = rays.view();
        check(rt_trace_ex(s_, &v, &r, (uint32_t)n, RT_ANY, opts, nullptr, bits.data(), nullptr,
                          nullptr),
              "rt_trace_ex (any)");
    }

    // Any hit with occluder hints (rt_trace_occluded_hinted): ray r first tries
    // triangle hintPrim[r % period] (period 0 = every ray its own slot; RT_MISS = no
    // hint), and with waveHints the hints of the other rays of its wave; the bits equal
    // occluded() for any hints.  Shadow rays laid out light by light take the primary
    // hits of the frame as hints with period = pixels.  The triangle-id -> leaf-record
    // map (which a refit leaves as it is) and the scratch are kept with the object.
    void occludedHinted(const Rays& rays, const DeviceArray<uint32_t>& hintPrim,
                        uint32_t period, DeviceArray<uint32_t>& bits, bool waveHints = true,
                        const rt_trace_opts* opts = nullptr) {
        const size_t n = rays.size();
        const size_t slots = period ? period : n;
        if (hintPrim.size() < slots) throw Error("occludedHinted: fewer hints than the period");
        bits.resize(n ? (n + 31) / 32 : 1);
        const rt_bvh_dev v = view();
        if (records_.size() != (size_t)std::max<uint32_t>(triCount_, 1u)) {
            records_.resize(std::max<uint32_t>(triCount_, 1u));
            check(rt_prim_records(s_, &v, triCount_, records_.data()), "rt_prim_records");
        }
        const rt_occluder_hint h = {hintPrim.data(), records_.data(), period, triCount_,
                                    waveHints ? RT_HINT_WAVE : 0u};
        size_t bytes = 0;
        check(rt_trace_occluded_hinted_scratch_bytes((uint32_t)n, opts, &h, &bytes),
              "rt_trace_occluded_hinted_scratch_bytes");
        if (hintScratch_.size() < bytes) {
            hip_check(hipStreamSynchronize(s_), "hipStreamSynchronize");  // old scratch in use
            hintScratch_.resize(bytes);
        }
        const rt_rays_soa r = rays.view();
        check(rt_trace_occluded_hinted(s_, &v, &r, (uint32_t)n, opts, &h, bits.data(),
                                       hintScratch_.data(), hintScratch_.size()),
              "rt_trace_occluded_hinted");
    }

    // Same topology, new positions (a deforming mesh with the same indices).
    template <class Mesh>
    void refit(const Mesh& mesh) {
        const std::vector<float> pos = packed_positions(mesh);
        pos_.upload(pos.data(), pos.size(), s_);
        if (idx_.size() == 0) idx_.upload(mesh.indices.data(), mesh.indices.size(), s_);
        DeviceArray<uint8_t> scratch(std::max<size_t>(256, (size_t)nodeCount_ * 4));
        check(rt_gpu_refit_bvh4(s_, pos_.data(), idx_.data(), nodes_.data(), nodeCount_,
                                tris_.data(), depth_, scratch.data(), scratch.size()),
              "rt_gpu_refit_bvh4");
        hip_check(hipStreamSynchronize(s_), "hipStreamSynchronize");  // scratch freed next
    }

    rt_bvh_dev view() const {
        return {nodes_.data(), tris_.data(), nodeCount_, triCount_, depth_, 0u};
    }
    // The node array as GpuBvh4Node bytes (downloads and waits).
    std::vector<rt_node4> downloadNodes() const { return nodes_.download(s_, nodeCount_); }
    std::vector<rt_tri64> downloadTriangles() const { return tris_.download(s_, triCount_); }
    uint32_t nodeCount() const { return nodeCount_; }
    uint32_t triCount() const { return triCount_; }
    uint32_t depth() const { return depth_; }
    hipStream_t stream() const { return s_; }

private:
    void build_device(uint32_t n, uint32_t maxp, uint32_t algorithm, uint32_t rounds) {
        rt_build_cfg cfg = {algorithm, maxp, 0u, rounds};
        uint32_t cap = 0;
        size_t scratchBytes = 0;
        check(rt_build_requirements(n, &cfg, &cap, &scratchBytes), "rt_build_requirements");
        nodes_.resize(cap ? cap : 1);
        tris_.resize(n ? n : 1);
        DeviceArray<rt_build_info> info(1);
        DeviceArray<uint8_t> scratch(scratchBytes ? scratchBytes : 256);
        check(rt_gpu_build(s_, pos_.data(), idx_.data(), n, &cfg, nodes_.data(), cap, tris_.data(),
                           info.data(), nullptr, nullptr, scratch.data(), scratch.size()),
              "rt_gpu_build");
        const std::vector<rt_build_info> hi = info.download(s_);  // waits: scratch freed next
        nodeCount_ = hi[0].node_count;
        depth_ = hi[0].depth;
        triCount_ = n;
    }

    hipStream_t s_ = nullptr;
    DeviceArray<float> pos_;
    DeviceArray<uint32_t> idx_;
    DeviceArray<rt_node4> nodes_;
    DeviceArray<rt_tri64> tris_;
    DeviceArray<uint32_t> records_;     // triangle id -> leaf record (occludedHinted)
    DeviceArray<uint8_t> hintScratch_;  // open masks and queue counters (occludedHinted)
    uint32_t nodeCount_ = 0, triCount_ = 0, depth_ = 0;
};

// A host-built SoftwareBVH (software_bvh.hpp:11-22,35-175) traversed on the device with
// the BVH2 loop of software_bvh_procedural.comp (rt_trace_bvh2): its GpuBvhNode and
// GpuTriangle arrays upload unchanged.  leafOrder (optional): the original triangle index
// of each leaf position, reported as the hit id (without it, the leaf position).
class GpuBVH2 {
public:
    static GpuBVH2 fromHost(const void* nodes, uint32_t nodeCount, const float* tris48,
                            const uint32_t* leafOrder, uint32_t triCount,
                            hipStream_t s = nullptr) {
        GpuBVH2 b;
        b.s_ = s;
        check(rt_bvh2_depth(nodes, nodeCount, &b.depth_), "rt_bvh2_depth");
        b.nodes_.upload(static_cast<const uint8_t*>(nodes), (size_t)nodeCount * 32, s);
        b.tris_.upload(tris48, (size_t)triCount * 12, s);
        if (leafOrder) b.order_.upload(leafOrder, triCount, s);
        b.nodeCount_ = nodeCount;
        b.triCount_ = triCount;
        return b;
    }

    void intersect(const Rays& rays, Hits& out, bool counters = false,
                   const rt_trace_opts* opts = nullptr) const {
        const size_t n = rays.size();
        out.prim.resize(n);
        out.t.resize(n);
        if (counters) {
            out.boxTests.resize(n);
            out.triTests.resize(n);
        }
        const rt_bvh2_dev v = view();
        const rt_rays_soa r = rays.view();
        rt_hits_soa h = {out.prim.data(), out.t.data()};
        check(rt_trace_bvh2(s_, &v, &r, (uint32_t)n, RT_CLOSEST | (counters ? RT_COUNTERS : 0u),
                            opts, &h, nullptr, counters ? out.boxTests.data() : nullptr,
                            counters ? out.triTests.data() : nullptr),
              "rt_trace_bvh2 (closest)");
    }

    void occluded(const Rays& rays, DeviceArray<uint32_t>& bits,
                  const rt_trace_opts* opts = nullptr) const {
        const size_t n = rays.size();
        bits.resize(n ? (n + 31) / 32 : 1);
        const rt_bvh2_dev v = view();
        const rt_rays_soa r = rays.view();
        check(rt_trace_bvh2(s_, &v, &r, (uint32_t)n, RT_ANY, opts, nullptr, bits.data(),
                            nullptr, nullptr),
              "rt_trace_bvh2 (any)");
    }

    rt_bvh2_dev view() const {
        return {nodes_.data(), tris_.data(), order_.size() ? order_.data() : nullptr,
                nodeCount_, triCount_, depth_, 0u};
    }
    uint32_t depth() const { return depth_; }

private:
    hipStream_t s_ = nullptr;
    DeviceArray<uint8_t> nodes_;
    DeviceArray<float> tris_;
    DeviceArray<uint32_t> order_;
    uint32_t nodeCount_ = 0, triCount_ = 0, depth_ = 0;
};

// The RADV acceleration-structure image of fast_bvh_builder.comp built on the device
// (rt_gpu_build_radv, at most 4,096 triangles as the reference builder) and traced in
// software over its bytes (rt_trace_radv): the ray-query consumers of that image.
class RadvBVH {
public:
    // order: RT_RADV_INPUT_ORDER (the reference builder's leaves) or RT_RADV_MORTON.
    template <class Mesh>
    static RadvBVH build(const Mesh& mesh, uint32_t order = RT_RADV_INPUT_ORDER,
                         hipStream_t s = nullptr) {
        const std::vector<float> pos = packed_positions(mesh);
        const uint32_t n = (uint32_t)(mesh.indices.size() / 3);
        RadvBVH b;
        b.s_ = s;
        DeviceArray<float> dpos;
        DeviceArray<uint32_t> didx;
        dpos.upload(pos.data(), pos.size(), s);
        didx.upload(mesh.indices.data(), mesh.indices.size(), s);
        uint32_t counts[32], offsets[32], levels = 0, boxes = 0;
        uint64_t bytes = 0;
        check(rt_tree_layout(n, counts, offsets, &levels, &boxes, &bytes), "rt_tree_layout");
        b.image_.resize(bytes);
        check(rt_gpu_build_radv(s, dpos.data(), didx.data(), n, order, b.image_.data(), bytes,
                                nullptr),
              "rt_gpu_build_radv");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");  // inputs freed next
        b.info_ = b.describe();
        b.depth_ = levels;
        return b;
    }

    // The header of the image (downloads and decodes it; rt_radv_describe).
    rt_radv_info describe() const {
        const std::vector<uint8_t> h = image_.download(s_);
        rt_radv_info info;
        check(rt_radv_describe(h.data(), h.size(), &info), "rt_radv_describe");
        return info;
    }

    void intersect(const Rays& rays, Hits& out, bool counters = false,
                   const rt_trace_opts* opts = nullptr) const {
        const size_t n = rays.size();
        out.prim.resize(n);
        out.t.resize(n);
        if (counters) {
            out.boxTests.resize(n);
            out.triTests.resize(n);
        }
        const rt_radv_dev v = view();
        const rt_rays_soa r = rays.view();
        rt_hits_soa h = {out.prim.data(), out.t.data()};
        check(rt_trace_radv(s_, &v, &r, (uint32_t)n, RT_CLOSEST | (counters ? RT_COUNTERS : 0u),
                            opts, &h, nullptr, counters ? out.boxTests.data() : nullptr,
                            counters ? out.triTests.data() : nullptr),
              "rt_trace_radv (closest)");
    }

    void occluded(const Rays& rays, DeviceArray<uint32_t>& bits,
                  const rt_trace_opts* opts = nullptr) const {
        const size_t n = rays.size();
        bits.resize(n ? (n + 31) / 32 : 1);
        const rt_radv_dev v = view();
        const rt_rays_soa r = rays.view();
        check(rt_trace_radv(s_, &v, &r, (uint32_t)n, RT_ANY, opts, nullptr, bits.data(), nullptr,
                            nullptr),
              "rt_trace_radv (any)");
    }

    rt_radv_dev view() const {
        return {image_.data(), image_.size(), info_.bvh_offset, info_.box_count,
                (uint32_t)info_.primitive_count, depth_};
    }
    const DeviceArray<uint8_t>& image() const { return image_; }

private:
    hipStream_t s_ = nullptr;
    DeviceArray<uint8_t> image_;
    rt_radv_info info_{};
    uint32_t depth_ = 0;
};

// ImageQualityMetrics of common/image_metrics.hpp:8-12.
struct ImageQuality {
    double mse = 0.0;
    double psnr = 100.0;
    double ssim = 1.0;
};

// ImageMetrics::compute(imgA, imgB, width, height) (image_metrics.hpp:16-104) of two
// float images in device memory (rt_image_metrics); waits for the result.
inline ImageQuality imageMetrics(const float* d_a, const float* d_b, uint32_t width,
                                 uint32_t height, hipStream_t s = nullptr) {
    size_t bytes = 0;
    check(rt_image_metrics_scratch_bytes(width, height, &bytes), "rt_image_metrics_scratch_bytes");
    DeviceArray<double> scratch(bytes / sizeof(double) + 1), out(3);
    check(rt_image_metrics(s, d_a, d_b, width, height, out.data(), nullptr, scratch.data(),
                           scratch.size() * sizeof(double)),
          "rt_image_metrics");
    const std::vector<double> h = out.download(s);  // waits: scratch freed next
    return {h[0], h[1], h[2]};
}

}  // namespace rtgpu
