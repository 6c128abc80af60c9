/*
 * oracle.h -- CPU restatement of the reference BVH build-and-traverse path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product (librtbvh.so) never links, calls or falls back to it.
 *
 * Every function cites the reference file:line it restates.  Arithmetic is
 * compiled with -ffp-contract=off so each float op rounds exactly once, in
 * the order written; the HIP kernels use the same order and the same flag,
 * so traversal t values and Morton codes agree bit for bit.
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   or_terrain, or_bvh2_build, or_bvh4_collapse, or_tree_layout,
 *   or_dynamic_grid        -> byte-compared against the reference headers
 *                             compiled in this container (oracle/_ref) and the
 *                             committed fixtures written from them.
 *   or_trace_bvh4          -> exp0_ray_stats.csv mean box/tri tests and hit
 *                             rates; known-answer rays (probe_main.cpp,
 *                             src/main.cpp); brute force on small meshes.
 *   or_exp0_rays           -> pinned through the same exp0 statistics.
 *   or_trace_bvh2          -> brute force over all triangles with the same
 *                             predicates; a hit is a hit of the BVH4
 *                             traversal of the same mesh up to the two
 *                             shaders' predicate differences (|a| 1e-8 vs 1e-7,
 *                             t > tmin vs t >= tmin); known-answer rays.
 *   or_morton30            -> bit-interleave closed form, per-axis brute force.
 *   or_sort_pairs          -> std::stable_sort on the same keys.
 *   or_lbvh_build          -> invariants: permutation, sortedness, parent box
 *                             = union of children, root = scene bounds,
 *                             layout == compute_tree_layout (_ref).
 *   or_lbvh_refit,
 *   or_bvh4_refit          -> on unchanged positions each reproduces the
 *                             builder's own tree bytes (implicit, radix,
 *                             software_bvh); after a deformation the two agree
 *                             on implicit trees and every slot box equals the
 *                             brute-force union of its records / child slots.
 *   or_radv_encode         -> static_assert offsets of radv_bvh.hpp:106-152,
 *                             header constants of fast_bvh_builder.comp:220-256,
 *                             and the N+1-ray verification of src/main.cpp.
 *   generators             -> closed forms (vertex on sphere radius bounds,
 *                             unit directions, hemisphere orientation).
 *   or_blur7               -> constant image is a fixed point; weights sum.
 *   or_bvh4_quantize,
 *   or_bvh4_dequantize     -> every decoded box contains its slot box; each
 *                             code is optimal (one step tighter breaks the
 *                             containment) and each exponent least, checked
 *                             by brute force over all 256 codes / exponents.
 */
#ifndef RT_ORACLE_H
#define RT_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* software_bvh.hpp:11-16 GpuBvhNode (32 B) */
typedef struct {
    float aabbMin[3];
    int32_t leftChild;
    float aabbMax[3];
    int32_t rightChild;
} or_bvh2_node;

/* software_bvh.hpp:178-187 GpuBvh4Node (128 B) */
typedef struct {
    float minX[4], maxX[4], minY[4], maxY[4], minZ[4], maxZ[4];
    int32_t child[4];
    int32_t count[4];
} or_bvh4_node;

/* radv_bvh.hpp:142-152 RadvBvhTriangleNode (64 B) */
typedef struct {
    float v[9];
    uint32_t reserved[3];
    uint32_t triangle_id;
    uint32_t geometry_id_and_flags;
    uint32_t reserved2;
    uint32_t id;
} or_tri64;

/* ---- mesh generators ---------------------------------------------------- */
uint32_t or_terrain_res(uint32_t target_tris);
void or_terrain(uint32_t target_tris, float size, float* pos, uint32_t* idx);
void or_terrain_ex(uint32_t target_tris, float size, int fused, float* pos, uint32_t* idx);
void or_dynamic_grid(uint32_t n, float t, float extent, float* xyz9, float* centroids);
void or_wave_grid(uint32_t nx, uint32_t ny, float extent, float t, float amp,
                  float fx, float fy, float* pos, uint32_t* idx);
void or_sphere(uint32_t nu, uint32_t nv, float radius, float disp, uint32_t seed,
               float* pos, uint32_t* idx);
void or_heightfield(uint32_t nx, uint32_t ny, float extent, float amp, uint32_t seed,
                    float* pos, uint32_t* idx);

/* ---- software_bvh builder ---------------------------------------------- */
int or_bvh2_build(const float* pos, uint32_t nverts, const uint32_t* idx, uint32_t ntris,
                  uint32_t max_prims, or_bvh2_node* nodes, uint32_t node_cap,
                  uint32_t* node_count, float* tris48, uint32_t* order);
int or_bvh4_collapse(const or_bvh2_node* n2, uint32_t count2, or_bvh4_node* n4,
                     uint32_t cap4, uint32_t* count4);
void or_tris48_to_tri64(const float* tris48, const uint32_t* order, uint32_t n, or_tri64* out);
uint32_t or_bvh4_depth(const or_bvh4_node* nodes, uint32_t count);

/* ---- quantized nodes (rt_qnode4; no reference counterpart) --------------
 * or_bvh4_quantize: per node and axis, e = least exponent in [-126, 127] with
 * dec(255) >= the max bound of the occupied slots (dec(q) = p + q * 2^e in
 * float, p = the min bound), qlo = greatest code with dec(q) <= the slot's
 * min, qhi = least code with dec(q) >= its max (binary searches over the
 * monotone dec).  or_bvh4_dequantize: the rt_node4 tree with the decoded
 * boxes (the tree a quantized trace is the exact trace of).                */
typedef struct {
    float p[3];
    int8_t e[3];
    uint8_t pad0;
    uint8_t qlo[3][4];
    uint8_t qhi[3][4];
    int16_t count[4];
    int32_t child[4];
} or_qnode4;
void or_bvh4_quantize(const or_bvh4_node* nodes, uint32_t count, or_qnode4* out);
void or_bvh4_dequantize(const or_qnode4* q, uint32_t count, or_bvh4_node* out);

/* ---- traversal ---------------------------------------------------------- */
enum { OR_CLOSEST = 0, OR_ANY = 1 };
void or_trace_bvh4(const or_bvh4_node* nodes, uint32_t node_count, const or_tri64* tris,
                   const float* ox, const float* oy, const float* oz, const float* tmin,
                   const float* dx, const float* dy, const float* dz, const float* tmax,
                   uint32_t nrays, int mode, uint32_t stack_limit, int nthreads,
                   uint32_t* out_prim, float* out_t, uint32_t* out_box, uint32_t* out_tri);
void or_trace_bvh2(const or_bvh2_node* nodes, uint32_t node_count, const float* tris48,
                   const uint32_t* order, const float* ox, const float* oy, const float* oz,
                   const float* tmin, const float* dx, const float* dy, const float* dz,
                   const float* tmax, uint32_t nrays, int mode, int nthreads,
                   uint32_t* out_prim, float* out_t, uint32_t* out_box, uint32_t* out_tri);
void or_trace_brute(const or_tri64* tris, uint32_t ntris,
                    const float* ox, const float* oy, const float* oz, const float* tmin,
                    const float* dx, const float* dy, const float* dz, const float* tmax,
                    uint32_t nrays, int mode, int nthreads, uint32_t* out_prim, float* out_t);

/* ---- ray generators ----------------------------------------------------- */
uint32_t or_pcg_hash(uint32_t v);
uint32_t or_ao_hash(uint32_t x);
void or_exp0_rays(uint32_t count, int coherent, const float cam[3], const float look[3],
                  float* ox, float* oy, float* oz, float* dx, float* dy, float* dz);
void or_camera_rays(uint32_t w, uint32_t h, const float eye[3], const float look[3],
                    const float up[3], float vfov_deg, float tmin, float tmax,
                    float* ox, float* oy, float* oz, float* tn,
                    float* dx, float* dy, float* dz, float* tx);
void or_shadow_rays(const float* pos, const uint32_t* idx,
                    const float* pox, const float* poy, const float* poz,
                    const float* pdx, const float* pdy, const float* pdz,
                    const uint32_t* prim, const float* t, uint32_t npix,
                    const float* lights, uint32_t nlights, float eps,
                    float* ox, float* oy, float* oz, float* tn,
                    float* dx, float* dy, float* dz, float* tx);
void or_ao_rays(const float* pos, const uint32_t* idx,
                const float* pox, const float* poy, const float* poz,
                const float* pdx, const float* pdy, const float* pdz,
                const uint32_t* prim, const float* t, uint32_t w, uint32_t h,
                uint32_t frame, float offset, float tmin, float tmax,
                float* ox, float* oy, float* oz, float* tn,
                float* dx, float* dy, float* dz, float* tx);
/* A two-level scene instance as the device's rt_instance: mesh for normals,
 * row-major 3x4 world-to-object map, kind 0 general / 1 translate /
 * 2 identity. */
typedef struct or_instance {
    const float* pos;
    const uint32_t* idx;
    float world_to_object[12];
    uint32_t kind;
} or_instance;
void or_ao_rays_scene(const or_instance* insts, uint32_t count,
                      const float* pox, const float* poy, const float* poz,
                      const float* pdx, const float* pdy, const float* pdz,
                      const uint32_t* prim, const uint32_t* inst, const float* t, uint32_t w,
                      uint32_t h, uint32_t frame, float offset, float tmin, float tmax,
                      float* ox, float* oy, float* oz, float* tn,
                      float* dx, float* dy, float* dz, float* tx);
void or_random_rays(uint32_t count, uint32_t seed, const float lo[3], const float hi[3],
                    float tmax, float* ox, float* oy, float* oz, float* tn,
                    float* dx, float* dy, float* dz, float* tx);

/* ---- LBVH (Morton + radix + implicit Box4) ------------------------------ */
void or_centroid_bounds(const float* pos, const uint32_t* idx, uint32_t ntris, float bounds[6]);
void or_morton30(const float* pos, const uint32_t* idx, uint32_t ntris,
                 const float bounds[6], uint32_t* codes);
uint32_t or_expand_bits10(uint32_t v);
void or_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t n);
void or_tree_layout(uint32_t n, uint32_t* counts, uint32_t* offsets, uint32_t* levels,
                    uint32_t* total_boxes, uint64_t* total_as_size);
int or_lbvh_build(const float* pos, const uint32_t* idx, uint32_t ntris,
                  or_bvh4_node* nodes, uint32_t node_cap, uint32_t* node_count,
                  or_tri64* tris, uint32_t* sorted_ids, uint32_t* codes_sorted);
int or_rlbvh_build_opt(const float* pos, const uint32_t* idx, uint32_t n, uint32_t max_prims,
                       uint32_t rounds, or_bvh2_node* n2, uint32_t cap2, uint32_t* count2,
                       or_bvh4_node* n4, uint32_t cap4, uint32_t* count4, or_tri64* tris,
                       uint32_t* sorted_ids, uint32_t* codes_sorted);
int or_rlbvh_build(const float* pos, const uint32_t* idx, uint32_t n, uint32_t max_prims,
                   or_bvh2_node* n2, uint32_t cap2, uint32_t* count2, or_bvh4_node* n4,
                   uint32_t cap4, uint32_t* count4, or_tri64* tris, uint32_t* sorted_ids,
                   uint32_t* codes_sorted);
int or_lbvh_refit(const float* pos, const uint32_t* idx, uint32_t ntris,
                  or_bvh4_node* nodes, uint32_t node_count, or_tri64* tris);
int or_bvh4_refit(const float* pos, const uint32_t* idx, or_bvh4_node* nodes,
                  uint32_t node_count, or_tri64* tris, uint32_t ntris);
int or_radv_decode(const uint8_t* img, uint64_t bytes, or_bvh4_node* nodes, uint32_t cap,
                   uint32_t* node_count, or_tri64* tris, uint32_t tcap, uint32_t* tri_count);
int or_radv_encode(const float* xyz9, uint32_t n, const uint32_t* order,
                   uint8_t* out, uint64_t out_bytes);

/* ---- post / metrics ----------------------------------------------------- */
void or_ao_resolve(const uint32_t* occl_bits, const uint32_t* prim_primary,
                   uint32_t npix, float* ao);
void or_blur7(const float* in, const float* depth, uint32_t w, uint32_t h,
              float depth_thr, float* out);
double or_rms(const float* a, const float* b, uint64_t n);
float or_percentile(float* v, uint32_t n, float q);
void or_timer_stats(float* v, uint32_t n, float out[8]);
void or_image_metrics(const float* a, const float* b, uint32_t w, uint32_t h, double out[3],
                      double* block);

#ifdef __cplusplus
}
#endif
#endif
