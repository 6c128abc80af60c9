// ref_shim.cpp -- extern "C" entry points over the reference's own header-only
// CPU path, compiled from the headers where they lie under the reference
// checkout (common/software_bvh.hpp, common/mesh_gen.hpp,
// common/image_metrics.hpp, include/radv_bvh.hpp).  Test infrastructure: builds oracle/_ref/libref.so,
// which the tests use to pin the restatement in oracle.cpp.  Nothing here is
// copied from the reference; it only calls the reference's functions.
#include "software_bvh.hpp"
#include "radv_bvh.hpp"
#include "image_metrics.hpp"

#include <cstring>

static SceneMesh make_mesh(const float* pos, uint32_t nverts, const uint32_t* idx,
                           uint32_t ntris) {
    SceneMesh m;
    m.vertices.resize(nverts);
    for (uint32_t i = 0; i < nverts; ++i) {
        for (int c = 0; c < 3; ++c) m.vertices[i].pos[c] = pos[i * 3 + c];
        m.vertices[i].normal[0] = 0.0f;
        m.vertices[i].normal[1] = 1.0f;
        m.vertices[i].normal[2] = 0.0f;
    }
    m.indices.assign(idx, idx + (size_t)ntris * 3);
    return m;
}

extern "C" {

// MeshGenerator::createTessellatedTerrain (mesh_gen.hpp:29)
int ref_terrain(uint32_t target, float size, float* pos, uint32_t* idx, uint32_t* nverts,
                uint32_t* ntris) {
    SceneMesh m = MeshGenerator::createTessellatedTerrain(target, size);
    *nverts = (uint32_t)m.vertices.size();
    *ntris = m.triangleCount();
    if (pos)
        for (size_t i = 0; i < m.vertices.size(); ++i)
            for (int c = 0; c < 3; ++c) pos[i * 3 + c] = m.vertices[i].pos[c];
    if (idx) std::memcpy(idx, m.indices.data(), m.indices.size() * 4);
    return 0;
}

// SoftwareBVH::build (software_bvh.hpp:35) + SoftwareBVH4::fromBVH2 (:193)
int ref_build(const float* pos, uint32_t nverts, const uint32_t* idx, uint32_t ntris,
              uint32_t max_prims, void* nodes2, uint32_t cap2, uint32_t* count2,
              void* nodes4, uint32_t cap4, uint32_t* count4, float* tris48) {
    SceneMesh m = make_mesh(pos, nverts, idx, ntris);
    SoftwareBVH b2 = SoftwareBVH::build(m, max_prims);
    SoftwareBVH4 b4 = SoftwareBVH4::fromBVH2(b2);
    *count2 = (uint32_t)b2.nodes.size();
    *count4 = (uint32_t)b4.nodes.size();
    if (b2.nodes.size() > cap2 || b4.nodes.size() > cap4) return 1;
    std::memcpy(nodes2, b2.nodes.data(), b2.nodes.size() * sizeof(GpuBvhNode));
    if (!b4.nodes.empty())
        std::memcpy(nodes4, b4.nodes.data(), b4.nodes.size() * sizeof(GpuBvh4Node));
    std::memcpy(tris48, b2.triangles.data(), b2.triangles.size() * sizeof(GpuTriangle));
    return 0;
}

// rdna2::compute_tree_layout (radv_bvh.hpp:193) and to_push_constants (:177)
int ref_tree_layout(uint32_t n, uint32_t* counts, uint32_t* offsets, uint32_t* levels,
                    uint32_t* total_boxes, uint32_t* total_as, void* push_constants) {
    rdna2::FastBvhTreeLayout l = rdna2::compute_tree_layout(n);
    *levels = l.numLevels;
    *total_boxes = l.totalBoxes;
    *total_as = l.totalAsSize;
    for (uint32_t i = 0; i < l.numLevels; ++i) {
        counts[i] = l.levelCounts[i];
        offsets[i] = l.levelOffsets[i];
    }
    if (push_constants) {
        rdna2::FastBvhPushConstants pc = l.to_push_constants(0, 0);
        std::memcpy(push_constants, &pc, sizeof pc);
    }
    return 0;
}

// rdna2::generate_dynamic_grid_mesh (radv_bvh.hpp:224)
int ref_dynamic_grid(uint32_t n, float t, float extent, float* xyz9, float* cen) {
    rdna2::DynamicMesh m = rdna2::generate_dynamic_grid_mesh(n, t, extent);
    std::memcpy(xyz9, m.vertices.data(), m.vertices.size() * 4);
    for (size_t i = 0; i < m.centroids.size(); ++i) {
        cen[i * 3 + 0] = m.centroids[i].x;
        cen[i * 3 + 1] = m.centroids[i].y;
        cen[i * 3 + 2] = m.centroids[i].z;
    }
    return (int)(m.vertices.size() / 9);
}

// Struct sizes / offsets the reference pins with static_assert.
int ref_sizes(uint32_t* out) {
    out[0] = sizeof(GpuBvhNode);
    out[1] = sizeof(GpuTriangle);
    out[2] = sizeof(GpuBvh4Node);
    out[3] = sizeof(rdna2::RadvAccelStructHeader);
    out[4] = sizeof(rdna2::RadvBvhBox32Node);
    out[5] = sizeof(rdna2::RadvBvhTriangleNode);
    out[6] = sizeof(rdna2::FastBvhPushConstants);
    out[7] = offsetof(rdna2::RadvBvhTriangleNode, triangle_id);
    return 8;
}

// RadvAccelStructHeader (radv_bvh.hpp:59-82) read through the reference's
// own struct: out = bvh_offset, root_flags, update_dispatch_size[3],
// compacted_size, serialization_size, copy_dispatch_size[3], size,
// geometry_type, geometry_count, leaf_nodes_offset, primitive_count,
// instance_count, leaf_node_offsets_offset, build_flags, type; aabb = min xyz,
// max xyz; root = the root Box32 node's child ids.
int ref_radv_header(const uint8_t* img, uint64_t* out, float* aabb, uint32_t* root) {
    rdna2::RadvAccelStructHeader h;
    std::memcpy(&h, img, sizeof h);
    const uint64_t v[] = {h.bvh_offset, h.root_flags, h.update_dispatch_size[0],
                          h.update_dispatch_size[1], h.update_dispatch_size[2], h.compacted_size,
                          h.serialization_size, h.copy_dispatch_size[0], h.copy_dispatch_size[1],
                          h.copy_dispatch_size[2], h.size, h.geometry_type, h.geometry_count,
                          h.leaf_nodes_offset, h.primitive_count, h.instance_count,
                          h.leaf_node_offsets_offset, h.build_flags, h.type};
    std::memcpy(out, v, sizeof v);
    aabb[0] = h.aabb.min.x; aabb[1] = h.aabb.min.y; aabb[2] = h.aabb.min.z;
    aabb[3] = h.aabb.max.x; aabb[4] = h.aabb.max.y; aabb[5] = h.aabb.max.z;
    rdna2::RadvBvhBox32Node r;
    std::memcpy(&r, img + h.bvh_offset, sizeof r);
    for (int i = 0; i < 4; ++i) root[i] = r.children[i];
    return (int)(sizeof v / sizeof v[0]);
}

// ImageMetrics::compute (image_metrics.hpp:16): out = mse, psnr, ssim
int ref_image_metrics(const float* a, const float* b, uint32_t w, uint32_t h, double* out) {
    const size_t n = (size_t)w * h;
    std::vector<float> va(a, a + n), vb(b, b + n);
    const ImageQualityMetrics m = ImageMetrics::compute(va, vb, w, h);
    out[0] = m.mse;
    out[1] = m.psnr;
    out[2] = m.ssim;
    return 0;
}

}  // extern "C"
