/*
 * rt_bvh.h -- C ABI of the MI355X (gfx950) BVH build-and-traverse library
 * (librtbvh.so).  Plain pointers and sizes only; every device pointer is
 * caller-owned HBM, every call is asynchronous on the given HIP stream and
 * returns an rt_status (no exceptions cross this boundary).
 *
 * Reference interfaces replaced (paths relative to the reference checkout):
 *   common/software_bvh.hpp:11-22,178-187  node / triangle layouts  -> rt_node4, rt_tri64
 *   common/software_bvh.hpp:35,193          SoftwareBVH::build / SoftwareBVH4::fromBVH2
 *                                           -> rt_gpu_build (device build of the same
 *                                              Box4 node format); a host-built
 *                                              SoftwareBVH4 node array is byte-identical
 *                                              to rt_node4 and its GpuTriangle array
 *                                              converts with rt_convert_tris48
 *   shaders/fast_bvh_builder.comp:74-83,91  FastBvhPushConstants + vkCmdDispatch(1,1,1)
 *                                           -> rt_gpu_build / rt_gpu_build_radv
 *   include/radv_bvh.hpp:193-215            compute_tree_layout -> rt_tree_layout
 *   include/radv_bvh.hpp:311-390            dump_bvh_structure -> rt_radv_describe
 *   rayQueryEXT over the fast_bvh_builder image -> rt_trace_radv
 *   benchmarks/exp0_probe/shaders/software_bvh_procedural.comp:47-171
 *                                           BVH2 traversal -> rt_trace_bvh2
 *   common/image_metrics.hpp:16-104         ImageMetrics::compute -> rt_image_metrics
 *   benchmarks/exp0_probe/shaders/software_bvh_bvh4.comp:22-46,60-196
 *                                           SSBO traversal ABI -> rt_trace
 *   benchmarks/exp0_probe/shaders/software_bvh.comp:5-15
 *                                           RayInput / HitResult -> rt_rays_soa / rt_hits_soa
 *   shaders/rq_ao.comp, shaders/tiled_lighting_ra.comp:136-155
 *                                           inline any-hit ray queries -> rt_trace(RT_ANY)
 */
#ifndef RT_BVH_H
#define RT_BVH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rt_status {
    RT_OK = 0,
    RT_ERR_ARG = 1,       /* null pointer, size mismatch, bad flag */
    RT_ERR_CAPACITY = 2,  /* an output or scratch buffer is too small */
    RT_ERR_HIP = 3,       /* a HIP runtime call failed (see rt_last_error) */
    RT_ERR_NOMEM = 4,     /* the library could not allocate its own state */
    RT_ERR_UNSUPPORTED = 5
} rt_status;

typedef void* rt_stream; /* hipStream_t; NULL = the null stream */

/* Box4 node, byte layout of GpuBvh4Node (software_bvh.hpp:178-187): per-child
 * SoA bounds, child[k] / count[k]:
 *   count > 0  leaf range of `count` rt_tri64 records starting at ~child
 *   count = -1 inner node index child
 *   count = 0  empty slot                                                  */
typedef struct rt_node4 {
    float minX[4], maxX[4], minY[4], maxY[4], minZ[4], maxZ[4];
    int32_t child[4];
    int32_t count[4];
} rt_node4;

/* 64-B triangle leaf, byte layout of RadvBvhTriangleNode
 * (radv_bvh.hpp:142-152): v0,v1,v2, original triangle index in triangle_id,
 * geometry_id_and_flags = 0x80000000 (opaque), id = 9.                     */
typedef struct rt_tri64 {
    float v[9];
    uint32_t reserved[3];
    uint32_t triangle_id;
    uint32_t geometry_id_and_flags;
    uint32_t reserved2;
    uint32_t id;
} rt_tri64;

/* A device-resident BVH.  depth = node levels on the longest path (sizes the
 * per-lane traversal stack: at most 3 * depth + 1 entries are ever live);
 * trees deeper than 64 levels are rejected (RT_ERR_CAPACITY).  0 = unknown:
 * the stack is sized for 64 levels and, should the tree be deeper, pushes
 * beyond it are dropped (never written past the stack).  Device builds
 * record the true depth in rt_build_info.depth (a radix tree is at most 62
 * levels; treelet restructuring may deepen it).                            */
typedef struct rt_bvh_dev {
    const rt_node4* nodes;  /* rt_node4[node_count], or rt_qnode4[] (format) */
    const rt_tri64* tris;
    uint32_t node_count;
    uint32_t tri_count;
    uint32_t depth;
    uint32_t format;        /* RT_NODES_F32 (0) or RT_NODES_Q8 */
} rt_bvh_dev;

/* Node formats of rt_bvh_dev.format.  RT_NODES_Q8: `nodes` points at
 * rt_qnode4 records (rt_gpu_quantize of the tree), traced by rt_trace /
 * rt_trace_ex (the one-ray-per-thread schedule); other trace entry points
 * return RT_ERR_UNSUPPORTED for it.                                         */
#define RT_NODES_F32 0u
#define RT_NODES_Q8 1u

/* 64-B quantized Box4 node (no reference counterpart; the layout of
 * compressed wide BVHs, Ylitie, Karras & Laine 2017, for four children):
 * child k's box on axis a is [p_a + qlo[a][k] * 2^e_a, p_a + qhi[a][k] *
 * 2^e_a] (one float rounding: the product is exact), rounded outward from
 * the rt_node4 slot box it encodes, so it contains that box; child / count
 * as rt_node4 (count as int16).  A trace over an RT_NODES_Q8 tree is the
 * trace of the rt_node4 tree with these (decoded) boxes: the same
 * triangles, and a superset of node visits of the exact tree.             */
typedef struct rt_qnode4 {
    float p[3];         /* per-axis origin: the min of the occupied slots */
    int8_t e[3];        /* per-axis scale exponents, -126..127            */
    uint8_t pad0;
    uint8_t qlo[3][4];  /* [axis][child]                                  */
    uint8_t qhi[3][4];
    int16_t count[4];
    int32_t child[4];
} rt_qnode4;

/* Structure-of-arrays rays: 8 float arrays of n entries (32 B per ray). */
typedef struct rt_rays_soa {
    const float *ox, *oy, *oz, *tmin, *dx, *dy, *dz, *tmax;
} rt_rays_soa;

/* Writable view of the same (for the ray generators). */
typedef struct rt_rays_out {
    float *ox, *oy, *oz, *tmin, *dx, *dy, *dz, *tmax;
} rt_rays_out;

/* Closest-hit results: prim = original triangle index (0xFFFFFFFF = miss),
 * t = hit distance (the ray's tmax on a miss).                              */
typedef struct rt_hits_soa {
    uint32_t* prim;
    float* t;
} rt_hits_soa;

#define RT_MISS 0xFFFFFFFFu

/* rt_trace flags */
#define RT_CLOSEST 0x0u  /* closest hit in [tmin, tmax) */
#define RT_ANY 0x1u      /* occlusion: any hit in [tmin, tmax), first-hit terminate */
#define RT_COUNTERS 0x100u /* also write per-ray box / triangle test counts */
#define RT_TRACE_PERSISTENT 0x200u /* persistent waves refilling finished lanes;
                                      same results, different schedule */

/* ---- library ------------------------------------------------------------ */
const char* rt_version(void);
const char* rt_last_error(void); /* text of the last RT_ERR_HIP on this thread */
rt_status rt_device_sync(void);

/* ---- layout (include/radv_bvh.hpp:193-215 compute_tree_layout) ----------
 * counts/offsets need room for 32 levels.  total_as_size = 192 + 128 boxes
 * + 64 n, the byte size of the RADV image rt_gpu_build_radv writes.        */
rt_status rt_tree_layout(uint32_t n, uint32_t* counts, uint32_t* offsets, uint32_t* levels,
                         uint32_t* total_boxes, uint64_t* total_as_size);

/* ---- device build (replaces fast_bvh_builder.comp / SoftwareBVH::build) --
 * Input: positions (float3 per vertex) + indices (uint3 per triangle), or
 * d_idx = NULL for 9 floats per triangle (fast_bvh_builder.comp:120-122).
 * Triangles are sorted by the 30-bit Morton code of their centroid
 * (centroid (p0 + p1 + p2) / 3 as software_bvh.hpp:53, ties by triangle
 * index) and the n rt_tri64 leaves are written in that order.  The tree:
 *   RT_BUILD_IMPLICIT  the fast_bvh_builder layout (radv_bvh.hpp:193-215):
 *                      complete 4-ary tree, leaf boxes of 4 consecutive
 *                      triangles, root first; node count = total_boxes.
 *   RT_BUILD_RADIX     binary radix tree over the codes (ranges of at most
 *                      max_leaf_prims triangles are leaves) collapsed with
 *                      SoftwareBVH4::collapse's rule (software_bvh.hpp:202-291),
 *                      pre-order numbering; at most max(1, n - 1) nodes.
 *   RT_BUILD_AUTO      IMPLICIT for n <= 4096 (one fused workgroup), else RADIX.
 * d_info (device memory) receives the node count and depth of the tree.
 * d_sorted_ids / d_codes (optional) receive the sorted triangle indices and
 * codes.  rt_build_requirements gives the node capacity and scratch bytes.  */
#define RT_BUILD_AUTO 0u
#define RT_BUILD_RADIX 1u
#define RT_BUILD_IMPLICIT 2u

typedef struct rt_build_cfg {
    uint32_t algorithm;      /* RT_BUILD_AUTO / RT_BUILD_RADIX / RT_BUILD_IMPLICIT */
    uint32_t max_leaf_prims; /* RADIX leaf size, 1..16; 0 = 4 (maxPrimsPerLeaf) */
    uint32_t sort_bits;      /* digit width of the multi-launch Morton radix sort
                                (n > 4096): 4, 6, 8 or 10 bits = 8, 5, 4 or 3
                                passes; 0 = 8.  The output does not depend on it. */
    uint32_t treelet_rounds; /* RADIX only: rounds of treelet restructuring (0..16, 0 =
                                none; no reference counterpart): bottom-up, every inner
                                node's treelet of up to 7 subtrees is rebuilt as the
                                binary tree of least SAH cost over them (Karras & Aila
                                2013) before the Box4 collapse.  Leaf ranges and the
                                triangle order are unchanged; hits and occlusion bits
                                are the same as without it (up to which of two
                                triangles at exactly equal t a closest-hit reports). */
} rt_build_cfg;

typedef struct rt_build_info {
    uint32_t node_count;
    uint32_t depth;
    uint32_t reserved[2];
} rt_build_info;

rt_status rt_build_requirements(uint32_t n, const rt_build_cfg* cfg, uint32_t* node_capacity,
                                size_t* scratch_bytes);
rt_status rt_gpu_build(rt_stream stream, const float* d_pos, const uint32_t* d_idx, uint32_t n,
                       const rt_build_cfg* cfg, rt_node4* d_nodes, uint32_t node_capacity,
                       rt_tri64* d_tris, rt_build_info* d_info, uint32_t* d_sorted_ids,
                       uint32_t* d_codes, void* d_scratch, size_t scratch_bytes);
/* rt_gpu_build that allocates nothing: no reference counterpart; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, checks
 * and results as rt_gpu_build, but the scratch comes from a fixed device
 * array in the library (10 MiB, one per GPU, kept until the process ends), so
 * n is at most RT_NOALLOC_BUILD_MAX (RT_ERR_CAPACITY beyond).  Every call on
 * one GPU shares that array: calls that overlap on one GPU (other streams,
 * other host threads) race and must be serialised by the caller.  The first
 * call on a device looks up the array's address there.                      */
#define RT_NOALLOC_BUILD_MAX 65536u
rt_status rt_gpu_build_noalloc(rt_stream stream, const float* d_pos, const uint32_t* d_idx,
                               uint32_t n, const rt_build_cfg* cfg, rt_node4* d_nodes,
                               uint32_t node_capacity, rt_tri64* d_tris, rt_build_info* d_info,
                               uint32_t* d_sorted_ids, uint32_t* d_codes);

/* Refit of an RT_BUILD_IMPLICIT tree for a deforming mesh: same topology
 * (leaf order taken from d_tris[i].triangle_id), new positions.  Scratch as
 * rt_build_requirements reports for RT_BUILD_IMPLICIT.                      */
rt_status rt_gpu_refit(rt_stream stream, const float* d_pos, const uint32_t* d_idx, uint32_t n,
                       rt_node4* d_nodes, uint32_t node_count, rt_tri64* d_tris,
                       void* d_scratch, size_t scratch_bytes);

/* Refit of ANY Box4 tree (RT_BUILD_RADIX or RT_BUILD_IMPLICIT device builds,
 * an uploaded SoftwareBVH4): same topology, leaf records re-read from their
 * triangle_id at the new positions, every slot box recomputed bottom-up (a
 * leaf range: union of its records' boxes; an inner slot: union of the
 * child's occupied slots).  Inner children must have larger node indices
 * than their parents (pre-order or level order, as every builder here and
 * SoftwareBVH4::fromBVH2 number them).  node_count = the tree's node count
 * (rt_build_info.node_count, not the capacity: every node is visited),
 * depth = its node levels (rt_build_info.depth, at most 64; 0 = 64: a
 * deeper tree must pass its depth and is rejected, RT_ERR_ARG);
 * scratch >= 4 * node_count bytes.  No
 * reference counterpart (the reference rebuilds deforming meshes per frame,
 * src/main.cpp:871-885).                                                    */
rt_status rt_gpu_refit_bvh4(rt_stream stream, const float* d_pos, const uint32_t* d_idx,
                            rt_node4* d_nodes, uint32_t node_count, rt_tri64* d_tris,
                            uint32_t depth, void* d_scratch, size_t scratch_bytes);

/* The RADV acceleration-structure image fast_bvh_builder.comp writes
 * (shaders/fast_bvh_builder.comp:91-257, formats radv_bvh.hpp:22-152):
 * 192-B header with the pre-root links at 0xB4 / 0xBC, box32 nodes (child
 * IDs (offset >> 3) | type, 24-B child boxes, empty slot ID 0xFFFFFFFF with
 * a NaN box) in the compute_tree_layout order, then the 64-B triangle nodes;
 * total_as_size bytes (rt_tree_layout).  n <= 4096, the reference builder's
 * limit (one workgroup there; here the fused path).  order:
 *   RT_RADV_INPUT_ORDER  leaves of 4 consecutive input triangles, as the
 *                        reference builder (d_sorted_ids, if given, = 0..n-1)
 *   RT_RADV_MORTON       the Morton order of rt_gpu_build.
 * d_image must be 64-byte aligned.                                          */
#define RT_RADV_INPUT_ORDER 0u
#define RT_RADV_MORTON 1u
rt_status rt_gpu_build_radv(rt_stream stream, const float* d_pos, const uint32_t* d_idx,
                            uint32_t n, uint32_t order, void* d_image, uint64_t image_bytes,
                            uint32_t* d_sorted_ids);

/* Decoded header of a RADV image in host memory: the RadvAccelStructHeader
 * fields (include/radv_bvh.hpp:59-82, the header part of dump_bvh_structure,
 * radv_bvh.hpp:311-390), the box32 node count (leaf_nodes_offset -
 * bvh_offset) / 128 and the root node's child ids.  RT_ERR_ARG when bytes <
 * 136 or the sections do not fit (bvh_offset not 64-aligned, the box section
 * not a whole number of 128-B nodes, triangles past the end); the header
 * fields are filled either way.                                             */
typedef struct rt_radv_info {
    uint32_t bvh_offset, root_flags;
    float aabb[6]; /* min xyz, max xyz */
    uint32_t update_dispatch_size[3];
    uint32_t geometry_type, geometry_count;
    uint32_t leaf_node_offsets_offset, build_flags, type;
    uint64_t compacted_size, serialization_size, size;
    uint32_t copy_dispatch_size[3];
    uint32_t box_count;
    uint64_t leaf_nodes_offset, primitive_count, instance_count;
    uint32_t root_children[4];
} rt_radv_info;
rt_status rt_radv_describe(const void* host_image, uint64_t bytes, rt_radv_info* info);


/* Batched build: many meshes of 1..RT_BATCH_BUILD_MAX triangles built in ONE
 * persistent launch, each with exactly the output rt_gpu_build gives it alone
 * with RT_BUILD_IMPLICIT (the fast_bvh_builder layout).  A mesh of at most
 * 4096 triangles is built in LDS (the fused path); a larger one by one
 * workgroup over its item scratch (rt_build_requirements' scratch bytes for
 * RT_BUILD_IMPLICIT; NULL / 0 for n <= 4096).  No reference counterpart
 * (fast_bvh_builder.comp builds one mesh per dispatch, src/main.cpp:871-885).
 *   1. rt_gpu_build_batch_plan (host) checks the items and writes the device
 *      table image (rt_gpu_build_batch_table_bytes(count) bytes) into host
 *      memory; order the items largest first for balance;
 *   2. the caller copies the table to device memory;
 *   3. rt_gpu_build_batch launches over it (no host work, no allocation:
 *      capturable in a graph).  workgroups = 0: min(count, 256).
 *      workgroups | RT_BUILD_BATCH_SHARED: the meshes above 4096 triangles
 *      are built by 256-thread workgroups with a few KB of LDS (a second
 *      launch), which fit on a compute unit beside other kernels' waves;
 *      the fused path's 1024-thread workgroup needs an idle compute unit.
 *      Same outputs; for builds issued while traces run (a serving loop).
 * Items must not share outputs or scratch.                                  */
#define RT_BATCH_BUILD_MAX 65536u
#define RT_BUILD_BATCH_SHARED 0x80000000u
typedef struct rt_build_item {
    const float* pos;       /* as rt_gpu_build */
    const uint32_t* idx;    /* NULL: 9 floats per triangle */
    rt_node4* nodes;        /* node_capacity >= rt_build_requirements' */
    rt_tri64* tris;
    rt_build_info* info;
    uint32_t* sorted_ids;   /* optional */
    uint32_t* codes;        /* optional */
    uint32_t n;
    uint32_t node_capacity;
    void* scratch;          /* n > 4096: the RT_BUILD_IMPLICIT scratch of n */
    uint64_t scratch_bytes;
} rt_build_item;
rt_status rt_gpu_build_batch_table_bytes(uint32_t count, size_t* bytes);
rt_status rt_gpu_build_batch_plan(const rt_build_item* items, uint32_t count, void* table,
                                  size_t table_bytes);
rt_status rt_gpu_build_batch(rt_stream stream, const void* d_table, uint32_t count,
                             uint32_t workgroups);

/* Upload-and-launch form of the batched build (plan, staging copy, launch in
 * one call; no allocation, no host wait), for a serving loop's deform-and-
 * rebuild requests.  gen (optional, one per item): the deform step, item i's
 * pos receives generate_dynamic_grid_mesh(items[i].n, t, extent) (9 floats
 * per triangle, include/radv_bvh.hpp:224-283, as rt_gen_dynamic_grid) in one
 * launch before the build; items[i].idx must then be NULL.  Staging and
 * device table as rt_trace_batch_submit (rt_gpu_build_batch_submit_bytes).  */
typedef struct rt_grid_gen {
    float t;
    float extent;
} rt_grid_gen;
rt_status rt_gpu_build_batch_submit_bytes(uint32_t count, int with_gen, size_t* bytes);
rt_status rt_gpu_build_batch_submit(rt_stream stream, const rt_build_item* items,
                                    const rt_grid_gen* gen, uint32_t count, uint32_t workgroups,
                                    void* pinned, size_t pinned_bytes, void* d_table,
                                    size_t table_bytes);

/* The RT_NODES_Q8 copy of a Box4 tree: qnodes[i] encodes nodes[i] (node
 * indices unchanged, same leaves), one thread per node.  Per axis, e is the
 * least exponent with p + 255 * 2^e >= the node's max bound (float
 * arithmetic), qlo the greatest code whose decoded bound is <= the slot's
 * min, qhi the least whose decoded bound is >= its max.  Leaf counts must
 * fit int16.  node_count = the tree's node count.  No reference counterpart.
 */
rt_status rt_gpu_quantize(rt_stream stream, const rt_node4* d_nodes, uint32_t node_count,
                          rt_qnode4* d_qnodes);

/* GpuTriangle records (48 B, software_bvh.hpp:18-22, leaf order) plus the
 * leaf-order original indices -> rt_tri64 leaves (software_bvh.hpp:66-84
 * drops the index; d_order restores it, NULL = identity).                  */
rt_status rt_convert_tris48(rt_stream stream, const float* d_tris48, const uint32_t* d_order,
                            uint32_t n, rt_tri64* d_out);

/* ---- trace (replaces software_bvh_bvh4.comp and the inline ray queries) --
 * RT_CLOSEST: hits->prim / hits->t per ray (same traversal order and
 *   predicates as software_bvh_bvh4.comp:111-190, so IDs and t are
 *   bit-identical to the CPU path on the same tree and rays).
 * RT_ANY: occl_bits bit r (word r/32, bit r%32) = ray r is occluded;
 *   occl_bits needs ceil(n/32) words.
 * RT_COUNTERS: box_tests[r] / tri_tests[r] (may be NULL otherwise).        */
rt_status rt_trace(rt_stream stream, const rt_bvh_dev* bvh, const rt_rays_soa* rays, uint32_t n,
                   uint32_t flags, rt_hits_soa* hits, uint32_t* occl_bits, uint32_t* box_tests,
                   uint32_t* tri_tests);

/* rt_trace with a schedule and launch-shape choice (same results for every
 * choice; the tuner in the Python host layer picks them by measurement):
 *   RT_SCHED_SIMPLE  one ray per thread, the hardware dispatcher balances.
 *   RT_SCHED_REFILL  each wave walks a chunk of rays_per_wave consecutive rays
 *                    and hands new rays to its finished lanes whenever at
 *                    least refill_lanes of 64 are idle.  With wave tiles
 *                    (occlusion bits only) the chunk is rays_per_wave / 64
 *                    consecutive tiles, refilled tile by tile.
 *   RT_SCHED_PACKET  (rt_trace_ex, RT_ANY without RT_COUNTERS or hit ids,
 *                    RT_NODES_F32 trees) a wave walks one node order for
 *                    its 64 rays: each stack entry carries the lanes whose
 *                    own box tests reached it, so every ray tests exactly the
 *                    triangles its own traversal reaches and its occlusion
 *                    bit is the same; node records are read once per wave.
 *                    For coherent rays (a light's shadow rays in wave tiles).
 *                    lds_stack does not apply; any_order must be 0.
 *   block            threads per workgroup: 64, 128 or 256 (0 = 256).
 *   lds_stack        traversal-stack entries per lane kept in LDS, 4..48
 *                    (0 = 20); deeper entries spill to private memory.  More
 *                    entries mean fewer spills and fewer resident waves.
 *   any_order        (rt_trace_ex, RT_ANY without RT_COUNTERS or hit ids,
 *                    RT_SCHED_SIMPLE) the order children are visited in:
 *                    RT_ORDER_NEAR (0, the reference's near-first order),
 *                    RT_ORDER_FAR (farthest hit child first), RT_ORDER_SLOT
 *                    (slot order, no sort).  The occlusion bits are the same
 *                    for every order (until an accepted hit ends a ray, every
 *                    node whose box meets [tmin, tmax] is visited); only the
 *                    work differs.  Other entry points take 0 only.
 *   grid_width,      (rt_trace_ex: RT_SCHED_SIMPLE, RT_SCHED_PACKET, and
 *   tile_width       RT_SCHED_REFILL for occlusion bits only) wave tiles:
 *                    the n rays are rows of grid_width consecutive rays (an
 *                    image, or several stacked: n a multiple of grid_width,
 *                    grid_width a multiple of tile_width); each wave traces
 *                    a tile of tile_width (8, 16, 32) rays across and
 *                    64 / tile_width rows down instead of 64 consecutive
 *                    rays.  Results are written per ray as before (same
 *                    values); only which rays share a wave changes.  0 / 0 =
 *                    linear waves.
 * opts = NULL: RT_SCHED_SIMPLE (RT_TRACE_PERSISTENT in flags: RT_SCHED_REFILL). */
#define RT_SCHED_SIMPLE 0u
#define RT_SCHED_REFILL 1u
#define RT_SCHED_PACKET 2u
#define RT_ORDER_NEAR 0u
#define RT_ORDER_FAR 1u
#define RT_ORDER_SLOT 2u
typedef struct rt_trace_opts {
    uint32_t schedule;
    uint32_t rays_per_wave; /* multiple of 64, at most 4096; 0 = 1024 */
    uint32_t refill_lanes;  /* 1..64; 0 = 16 */
    uint32_t block;         /* 64 / 128 / 256; 0 = 256 */
    uint32_t lds_stack;     /* 4..48; 0 = 20 */
    uint32_t any_order;     /* RT_ORDER_*; 0 = near first */
    uint32_t grid_width;    /* rays per row of a ray grid; 0 = linear waves */
    uint32_t tile_width;    /* 8 / 16 / 32 with grid_width */
} rt_trace_opts;
rt_status rt_trace_ex(rt_stream stream, const rt_bvh_dev* bvh, const rt_rays_soa* rays,
                      uint32_t n, uint32_t flags, const rt_trace_opts* opts, rt_hits_soa* hits,
                      uint32_t* occl_bits, uint32_t* box_tests, uint32_t* tri_tests);

/* Occlusion bits with occluder hints: the shadow ray queries of
 * benchmarks/expB_shadows/shaders/lighting_rayquery_tiled.comp:110-129 and
 * shaders/tiled_lighting_ra.comp (TerminateOnFirstHit queries from surface
 * points), given per ray a triangle likely to occlude it -- for shadow rays
 * the triangle the primary ray hit (a shadow cache: Haines & Greenberg 1986,
 * "The light buffer").  The bits equal rt_trace_ex(RT_ANY) on the same tree
 * and rays, for any hints: a hint settles a ray only when the triangle passes
 * the hit predicate in [tmin, tmax) AND its own box (min / max of its
 * vertices) passes the traversal's slab test, which implies that the
 * traversal reaches it (the slab test is monotone in the box and every slot
 * box contains the boxes of the triangles under it: true of every tree this
 * library builds or refits and of a SoftwareBVH4 upload); rays no hint
 * settles are traversed.
 *   prim    hint slot s: a triangle id (triangle_id of the leaf records), or
 *           RT_MISS / any id >= ids for none; ray r uses slot r % period
 *           (period 0 = n), e.g. the primary hits of a frame for its shadow
 *           rays laid out light by light (r = light * pixels + pixel).
 *   record  triangle id -> leaf record index, ids entries (rt_prim_records).
 *   flags   RT_HINT_WAVE: a ray also tries the hints of the other rays of its
 *           wave of 64 (wave tiles: its tile), one triangle for the wave at a
 *           time.
 * opts: block, lds_stack, grid_width / tile_width as rt_trace_ex; the rays
 * the hints leave open are traced with the refill schedule (refill_lanes) by
 * persistent waves that take segments of rays_per_wave ray slots (a
 * multiple of 64, at most 4096) one after another.  Scratch: device memory
 * of rt_trace_occluded_hinted_scratch_bytes bytes (256-B aligned), reusable
 * once the stream has passed the call.  RT_NODES_F32 trees only.           */
#define RT_HINT_WAVE 1u
typedef struct rt_occluder_hint {
    const uint32_t* prim;
    const uint32_t* record;
    uint32_t period;
    uint32_t ids;
    uint32_t flags;
} rt_occluder_hint;
rt_status rt_trace_occluded_hinted_scratch_bytes(uint32_t n, const rt_trace_opts* opts,
                                                 const rt_occluder_hint* hint, size_t* bytes);
rt_status rt_trace_occluded_hinted(rt_stream stream, const rt_bvh_dev* bvh,
                                   const rt_rays_soa* rays, uint32_t n, const rt_trace_opts* opts,
                                   const rt_occluder_hint* hint, uint32_t* occl_bits,
                                   void* d_scratch, size_t scratch_bytes);
/* record[id] = the leaf record holding triangle id (RT_MISS where none), for
 * ids ids, from the tree's leaf records (no reference counterpart: the
 * reference's GpuTriangle records carry no id).                            */
rt_status rt_prim_records(rt_stream stream, const rt_bvh_dev* bvh, uint32_t ids,
                          uint32_t* record);

/* Software traversal of a RADV image in device memory: the consumer the
 * reference's hardware ray queries are for (rayQueryEXT over the
 * fast_bvh_builder output, src/main.cpp:871-885 then the a17 shaders), with
 * rt_trace's predicates and near-first order over each box32 node's four
 * slots; a triangle child is a one-triangle leaf, its triangle_id the hit id.
 * On the image of a tree, results and counters equal rt_trace on the same
 * tree as Box4 nodes (rt_gpu_build RT_BUILD_IMPLICIT for the Morton image).
 * Children of other node types, and ids outside their section, are skipped.
 * image 16-byte aligned; box_count / tri_count from rt_radv_describe (or
 * rt_tree_layout for this library's images); depth = box levels (0 =
 * unknown, at most 64).  opts: block and lds_stack only.                    */
typedef struct rt_radv_dev {
    const void* image;
    uint64_t bytes;
    uint32_t bvh_offset;
    uint32_t box_count;
    uint32_t tri_count;
    uint32_t depth;
} rt_radv_dev;
rt_status rt_trace_radv(rt_stream stream, const rt_radv_dev* img, const rt_rays_soa* rays,
                        uint32_t n, uint32_t flags, const rt_trace_opts* opts, rt_hits_soa* hits,
                        uint32_t* occl_bits, uint32_t* box_tests, uint32_t* tri_tests);

/* BVH2 traversal of a SoftwareBVH as built (software_bvh.hpp:11-22,35-175:
 * GpuBvhNode 32 B nodes, GpuTriangle 48 B triangles in leaf order), the
 * loop and predicates of benchmarks/exp0_probe/shaders/
 * software_bvh_procedural.comp:47-171: popped node's box tested against the
 * closest t (tFar >= max(0, tNear), tNear < closest), leaf triangles in
 * record order (|a| < 1e-8 rejected, accept t > tmin and t < closest),
 * inner node: both children's boxes tested, hit ones pushed far first.
 * order (optional) maps a leaf position to the original triangle index (the
 * index SoftwareBVH drops); NULL reports the leaf position.  depth = node
 * levels (rt_bvh2_depth on the host array; 0 = unknown, at most 64).
 * flags / outputs as rt_trace; opts: block and lds_stack only.            */
typedef struct rt_bvh2_dev {
    const void* nodes;     /* GpuBvhNode[node_count] */
    const float* tris48;   /* GpuTriangle[tri_count] */
    const uint32_t* order; /* optional: original index of each leaf position */
    uint32_t node_count;
    uint32_t tri_count;
    uint32_t depth;
    uint32_t reserved;
} rt_bvh2_dev;
rt_status rt_bvh2_depth(const void* host_nodes, uint32_t node_count, uint32_t* depth);
rt_status rt_trace_bvh2(rt_stream stream, const rt_bvh2_dev* bvh, const rt_rays_soa* rays,
                        uint32_t n, uint32_t flags, const rt_trace_opts* opts, rt_hits_soa* hits,
                        uint32_t* occl_bits, uint32_t* box_tests, uint32_t* tri_tests);

/* Batched trace: many (tree, ray set, mode) items in ONE launch, each with
 * the results rt_trace gives it alone.  No reference counterpart (the
 * reference dispatches one ray set per submit, microbench_main.cpp:344);
 * this is the batching path over the same rt_trace contract.
 *   1. rt_trace_batch_plan (host): checks the items and fills
 *      chunk_start[0..count] (item i owns the 64-ray chunks
 *      [chunk_start[i], chunk_start[i+1])) and the deepest tree.
 *   2. the caller copies items[] and chunk_start[] to device memory;
 *   3. rt_trace_batch launches over those device copies (one kernel, no
 *      host work, no allocation: capturable in a graph).                    */
typedef struct rt_trace_item {
    rt_bvh_dev bvh;
    rt_rays_soa rays;
    uint32_t count;           /* rays */
    uint32_t flags;           /* RT_CLOSEST or RT_ANY */
    rt_hits_soa hits;         /* closest: prim and t; any: prim optional (NULL) */
    uint32_t* occlusion_bits; /* any: ceil(count / 32) words */
} rt_trace_item;
rt_status rt_trace_batch_plan(const rt_trace_item* items, uint32_t count, uint32_t* chunk_start,
                              uint32_t* depth);
rt_status rt_trace_batch(rt_stream stream, const rt_trace_item* d_items,
                         const uint32_t* d_chunk_start, uint32_t count, uint32_t total_chunks,
                         uint32_t depth);
/* rt_trace_batch with the schedule and launch shape of opts (same results
 * for every choice).  RT_SCHED_SIMPLE: d_chunk_start from rt_trace_batch_plan
 * (64-ray chunks, one per wave).  RT_SCHED_REFILL: d_chunk_start from
 * rt_trace_batch_plan_ex with chunk_rays = opts->rays_per_wave; each wave
 * walks one chunk, refilling finished lanes (refill_lanes as rt_trace_ex). */
rt_status rt_trace_batch_ex(rt_stream stream, const rt_trace_item* d_items,
                            const uint32_t* d_chunk_start, uint32_t count, uint32_t total_chunks,
                            uint32_t depth, const rt_trace_opts* opts);
/* rt_trace_batch_plan over chunks of chunk_rays rays (a multiple of 64, at
 * most 4096): item i owns chunks [chunk_start[i], chunk_start[i+1]).      */
rt_status rt_trace_batch_plan_ex(const rt_trace_item* items, uint32_t count, uint32_t chunk_rays,
                                 uint32_t* chunk_start, uint32_t* depth);

/* Upload-and-launch form of the batched trace, for a serving loop: ONE call
 * per batch of requests, no allocation, no host wait.  No reference
 * counterpart (the reference submits one dispatch per ray set and waits,
 * microbench_main.cpp:344, vk_timer.cpp:88-89).  From the host items (and,
 * if gen != NULL, one synthetic payload per item) it plans the chunk table
 * (chunks of opts->rays_per_wave rays for RT_SCHED_REFILL, else 64), writes
 * items, payloads and chunk table into the caller's pinned host staging
 * (hipHostMalloc'd, rt_trace_batch_submit_bytes bytes), copies them to the
 * caller's device table (same size, 256-B aligned) on the stream, generates
 * the payloads (gen: item i's rays are rt_gen_random_rays(count, seed, lo,
 * hi, tmax), written through items[i].rays) and launches rt_trace_batch_ex.
 * The staging may be reused once the stream has passed this call (the copy
 * reads it asynchronously).                                                  */
typedef struct rt_ray_gen {
    uint32_t seed;
    float lo[3], hi[3]; /* origins uniform in [lo, hi), directions on the sphere */
    float tmax;         /* tmin 0 */
} rt_ray_gen;
rt_status rt_trace_batch_submit_bytes(uint32_t count, int with_gen, size_t* bytes);
rt_status rt_trace_batch_submit(rt_stream stream, const rt_trace_item* items,
                                const rt_ray_gen* gen, uint32_t count, const rt_trace_opts* opts,
                                void* pinned, size_t pinned_bytes, void* d_table,
                                size_t table_bytes);

/* ---- two-level scene (the reference traces through a TLAS of instances:
 * src/main.cpp:920-1000, fullframe_benchmark.cpp:207-219) ------------------
 * A flat top level: every ray visits the instances in order, each instance a
 * BLAS placed by a world-to-object affine map.  Closest hit carries the best
 * t from instance to instance (a later instance must be strictly nearer, so
 * ties keep the lower instance); any hit stops at the first occluding
 * instance.  An instance's result equals rt_trace of its BLAS on the mapped
 * ray with tmax = the best t so far.  d_inst (closest, optional) receives the
 * instance index of the hit (RT_MISS on a miss).                            */
#define RT_INST_GENERAL 0u   /* o' = M o + T, d' = M d (M = 3x3 part)   */
#define RT_INST_TRANSLATE 1u /* M is the identity: o' = o + T, d' = d   */
#define RT_INST_IDENTITY 2u  /* world space is object space             */
typedef struct rt_instance {
    rt_bvh_dev blas;
    const float* pos;          /* the instance's mesh, for hit normals: vertices */
    const uint32_t* idx;       /* triangles; NULL = 9 floats per triangle in pos */
    float world_to_object[12]; /* row-major 3x4 [M | T] */
    uint32_t kind;             /* RT_INST_* */
    uint32_t reserved[7];
} rt_instance;
rt_status rt_trace_scene(rt_stream stream, const rt_instance* d_instances, uint32_t count,
                         uint32_t max_depth, const rt_rays_soa* rays, uint32_t n,
                         uint32_t flags, rt_hits_soa* hits, uint32_t* d_inst,
                         uint32_t* occl_bits);
/* rt_trace_scene with the launch shape of opts (block, lds_stack; schedule
 * must be RT_SCHED_SIMPLE); same results for every shape.                   */
rt_status rt_trace_scene_ex(rt_stream stream, const rt_instance* d_instances, uint32_t count,
                            uint32_t max_depth, const rt_rays_soa* rays, uint32_t n,
                            uint32_t flags, const rt_trace_opts* opts, rt_hits_soa* hits,
                            uint32_t* d_inst, uint32_t* occl_bits);

/* ---- frame post-processing (shaders/rq_ao.comp, shaders/bilateral_blur.comp) */
/* AO rays of rt_gen_ao_rays for a two-level scene: the hit normal comes from
 * instance d_inst[p]'s mesh (mapped to world space for RT_INST_GENERAL).    */
rt_status rt_gen_ao_rays_scene(rt_stream stream, const rt_instance* d_instances, uint32_t count,
                               const rt_rays_soa* primary, const uint32_t* d_prim,
                               const uint32_t* d_inst, const float* d_t, uint32_t w,
                               uint32_t h, uint32_t frame, float offset, float tmin, float tmax,
                               const rt_rays_out* out);
/* 1 SPP AO: ao[p] = 1 for a primary miss, else 0 if the AO ray is occluded. */
rt_status rt_ao_resolve(rt_stream stream, const uint32_t* d_bits, const uint32_t* d_prim,
                        uint32_t npix, float* d_ao);
/* 7x7 depth-gated bilateral filter: spatial weights w[|i|] w[|j|] from the
 * first four taps of bilateral_blur.comp:28-31, a neighbour counts iff it is
 * inside the image and |depth - centre depth| < depth_thr
 * (bilateral_blur.comp:41,49); out = sum / max(total, 0.001); pixels with
 * depth >= 1e30 (background) pass through (bilateral_blur.comp:23-26).     */
rt_status rt_blur7(rt_stream stream, const float* d_in, const float* d_depth, uint32_t w,
                   uint32_t h, float depth_thr, float* d_out);

/* ---- image quality (common/image_metrics.hpp:16-104 ImageMetrics::compute) --
 * d_out = {mse, psnr, ssim} (3 doubles, device memory) of two w x h float
 * images: mse = mean of (double(a) - double(b))^2; psnr = 100 when
 * mse < 1e-12, else 10 log10(1 / mse); ssim = mean over the full 8 x 8 blocks
 * of the block SSIM (double means, (n - 1)-normalised variances and
 * covariance, C1 = 0.01^2, C2 = 0.03^2) clamped to [0, 1], 1.0 when no full
 * block fits; no pixels: {0, 100, 1}.  d_block_ssim (optional): each block's
 * clamped SSIM, row-major over the (w / 8) x (h / 8) blocks, equal to the
 * reference's per-block value bit for bit.  The totals add the same terms as
 * the reference in another (fixed) order.  Scratch: rt_image_metrics_scratch_
 * bytes, 8-byte aligned; the call replaces ImageMetrics::compute's host
 * loops at fullframe_benchmark.cpp:1045 and the experiment mains. */
rt_status rt_image_metrics_scratch_bytes(uint32_t w, uint32_t h, size_t* bytes);
rt_status rt_image_metrics(rt_stream stream, const float* d_a, const float* d_b, uint32_t w,
                           uint32_t h, double* d_out, double* d_block_ssim, void* d_scratch,
                           size_t scratch_bytes);

/* ---- generators (synthetic inputs made on the device) -------------------- */
rt_status rt_gen_sphere(rt_stream stream, uint32_t nu, uint32_t nv, float radius, float disp,
                        uint32_t seed, float* d_pos, uint32_t* d_idx);
rt_status rt_gen_wave_grid(rt_stream stream, uint32_t nx, uint32_t ny, float extent, float t,
                           float amp, float fx, float fy, float* d_pos, uint32_t* d_idx);
rt_status rt_gen_heightfield(rt_stream stream, uint32_t nx, uint32_t ny, float extent, float amp,
                             uint32_t seed, float* d_pos, uint32_t* d_idx);
/* common/mesh_gen.hpp:29-83 MeshGenerator::createTessellatedTerrain(target,
 * size): res = floor(sqrt(target / 2)) + 1 (at least 2) vertices per side
 * (rt_terrain_resolution), 2 (res - 1)^2 triangles.  d_pos: 3 res^2 floats
 * (x, height, z); d_idx: 6 (res - 1)^2 indices.  fused != 0 evaluates the
 * three multiply-adds as the reference's -O3 -march=native build contracts
 * them (the build behind the reference's published results); heights use the device sinf /
 * cosf, within a few ulp of the host library's.                             */
rt_status rt_terrain_resolution(uint32_t target, uint32_t* res, uint32_t* nverts,
                                uint32_t* ntris);
rt_status rt_gen_terrain(rt_stream stream, uint32_t target, float size, int fused, float* d_pos,
                         uint32_t* d_idx);
/* include/radv_bvh.hpp:224-283 generate_dynamic_grid_mesh: 2 triangles per
 * quad, so d_xyz9 receives 2 * floor(n / 2) triangles (9 floats each); an
 * odd n leaves the last of n slots untouched.                              */
rt_status rt_gen_dynamic_grid(rt_stream stream, uint32_t n, float t, float extent, float* d_xyz9);
rt_status rt_gen_camera_rays(rt_stream stream, uint32_t w, uint32_t h, const float eye[3],
                             const float look[3], const float up[3], float vfov_deg, float tmin,
                             float tmax, const rt_rays_out* out);
rt_status rt_gen_shadow_rays(rt_stream stream, const float* d_pos, const uint32_t* d_idx,
                             const rt_rays_soa* primary, const uint32_t* d_prim,
                             const float* d_t, uint32_t npix, const float* d_lights,
                             uint32_t nlights, float eps, const rt_rays_out* out);
rt_status rt_gen_ao_rays(rt_stream stream, const float* d_pos, const uint32_t* d_idx,
                         const rt_rays_soa* primary, const uint32_t* d_prim, const float* d_t,
                         uint32_t w, uint32_t h, uint32_t frame, float offset, float tmin,
                         float tmax, const rt_rays_out* out);
/* benchmarks/exp0_probe/shaders/software_bvh_bvh4.comp:50-94 (the exp0
 * probe's procedural rays, microbench_main.cpp:290-294 passes cam (0,50,-45),
 * look (0,0,5)): coherent != 0, ray i is pixel (i % 1024, i / 1024) of a
 * 1024-wide grid, dir = normalize(fwd + 0.55 u right + 0.55 v up); else
 * origin (U*80-40, U*15+15, U*80-40) and dir normalize(U*0.8-0.4, -1-U*0.5,
 * U*0.8-0.4) from pcg_hash(6 i + k) * (1 / (2^32 - 1)).  Every ray gets tmin /
 * tmax; bit-identical to the shader's float arithmetic with one rounding per
 * operator.                                                               */
rt_status rt_gen_exp0_rays(rt_stream stream, uint32_t count, int coherent, const float cam[3],
                           const float look[3], float tmin, float tmax, const rt_rays_out* out);
rt_status rt_gen_random_rays(rt_stream stream, uint32_t count, uint32_t seed, const float lo[3],
                             const float hi[3], float tmax, const rt_rays_out* out);

/* Per-pixel light masks from light-major occlusion bits
 * (bit l*npix + p -> mask[p] bit l), nlights <= 16.                         */
rt_status rt_pack_light_masks(rt_stream stream, const uint32_t* d_bits, uint32_t npix,
                              uint32_t nlights, uint16_t* d_masks);

#ifdef __cplusplus
}
#endif
#endif /* RT_BVH_H */
