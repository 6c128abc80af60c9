// trace.hip -- Box4 BVH traversal on gfx950: closest-hit and any-hit
// (occlusion) over rt_node4 / rt_tri64 trees, SoA rays.
//
// Closest hit follows benchmarks/exp0_probe/shaders/software_bvh_bvh4.comp:
// 96-190: the same inverse-direction guard, slab test, child filter, child
// order (the shader's exchange sort by descending entry t, far children
// pushed first), leaf-first processing and Moller-Trumbore predicate, so on
// the same tree and rays the hit ID, t and box/triangle counters equal the
// CPU path bit for bit.  Any hit (the role of the TerminateOnFirstHit ray
// queries in rq_ao.comp / tiled_lighting_ra.comp) walks the same near-first
// order and stops at the first accepted triangle.
//
// Per-lane stack: the top `ls` entries live in LDS (entry-major, so lanes at
// the same depth hit distinct banks; the default 20 entries = 5 KB per wave
// keep 8 waves per SIMD resident), deeper entries go to a private array.
// The LDS depth and the block size are launch parameters (dynamic LDS), so a
// tuner can trade occupancy against private-memory spills; nothing is ever
// dropped (the shader drops on overflow, :185-187): 3 * depth + 1 entries
// bound the live stack and the private size is chosen from the tree depth.
#include <cstring>
#include <vector>

#include "rt_internal.h"

namespace rt {
namespace {

constexpr int kBlock = 256;         // largest block; launches use 64, 128 or 256
constexpr uint32_t kLdsStack = 20;  // default LDS entries per lane
constexpr uint32_t kLdsMin = 4, kLdsMax = 48;
constexpr int kPrivSmall = 48;      // private entries when 3 * depth + 1 - ls <= 48
constexpr int kPrivLarge = 3 * (int)kMaxDepth + 1 - (int)kLdsMin;

struct Child {
    float key;
    int32_t child;
    int32_t count;
};

// Order constants of an any-hit query that returns only occlusion bits
// (rt_trace_opts.any_order): near-first (the reference's), far-first, slot.
// kOrderInner (internal, the default for bits-only queries): inner children
// near-first, each placed on the stack at its rank (no sorted copy of the
// node), leaf children tested in slot order.
constexpr int kOrderNear = 0, kOrderFar = 1, kOrderSlot = 2, kOrderInner = 3;

// Compare-exchange putting the smaller key first (the far-first order's
// ascending sort; pushes then go nearest first, so the farthest pops first).
__device__ __forceinline__ void cx_up(Child& a, Child& b) {
    const bool s = b.key < a.key;
    const float ka = s ? b.key : a.key, kb = s ? a.key : b.key;
    const int32_t ca = s ? b.child : a.child, cb = s ? a.child : b.child;
    const int32_t na = s ? b.count : a.count, nb = s ? a.count : b.count;
    a = Child{ka, ca, na};
    b = Child{kb, cb, nb};
}

__device__ __forceinline__ void cx(Child& a, Child& b) {
    // software_bvh_bvh4.comp:151: swap when st[i] < st[j].  Written as
    // selects (one compare, two v_cndmask per field) rather than a branch:
    // a divergent branch costs exec-mask bookkeeping and register moves.
    const bool s = a.key < b.key;
    const float ka = s ? b.key : a.key, kb = s ? a.key : b.key;
    const int32_t ca = s ? b.child : a.child, cb = s ? a.child : b.child;
    const int32_t na = s ? b.count : a.count, nb = s ? a.count : b.count;
    a = Child{ka, ca, na};
    b = Child{kb, cb, nb};
}

// Global-address-space views.  Kernel pointer arguments are global to the
// compiler, but pointers read from memory (the batch items) are generic, and
// generic accesses compile to flat loads (which also wait on the LDS
// counter); the traversal is written over either pointer kind.
// (clang vector types, not HIP's float4 class, whose copy constructor takes
// a generic reference.)
#define RT_GLOBAL __attribute__((address_space(1)))
typedef float f4v __attribute__((ext_vector_type(4)));
typedef int32_t i4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ const f4v* f4(const rt_node4* p) {
    return reinterpret_cast<const f4v*>(p);
}
__device__ __forceinline__ const RT_GLOBAL f4v* f4(const RT_GLOBAL rt_node4* p) {
    return reinterpret_cast<const RT_GLOBAL f4v*>(p);
}
__device__ __forceinline__ const f4v* f4(const rt_tri64* p) {
    return reinterpret_cast<const f4v*>(p);
}
__device__ __forceinline__ const RT_GLOBAL f4v* f4(const RT_GLOBAL rt_tri64* p) {
    return reinterpret_cast<const RT_GLOBAL f4v*>(p);
}
// Float k of a triangle record.
__device__ __forceinline__ float f1(const rt_tri64* p, int k) {
    return reinterpret_cast<const float*>(p)[k];
}
__device__ __forceinline__ float f1(const RT_GLOBAL rt_tri64* p, int k) {
    return reinterpret_cast<const RT_GLOBAL float*>(p)[k];
}
__device__ __forceinline__ const char* bytes(const rt_node4* p) {
    return reinterpret_cast<const char*>(p);
}
__device__ __forceinline__ const RT_GLOBAL char* bytes(const RT_GLOBAL rt_node4* p) {
    return reinterpret_cast<const RT_GLOBAL char*>(p);
}
__device__ __forceinline__ const i4v* i4p(const char* p) {
    return reinterpret_cast<const i4v*>(p);
}
__device__ __forceinline__ const RT_GLOBAL i4v* i4p(const RT_GLOBAL char* p) {
    return reinterpret_cast<const RT_GLOBAL i4v*>(p);
}
__device__ __forceinline__ const f4v* f4p(const char* p) {
    return reinterpret_cast<const f4v*>(p);
}
__device__ __forceinline__ const RT_GLOBAL f4v* f4p(const RT_GLOBAL char* p) {
    return reinterpret_cast<const RT_GLOBAL f4v*>(p);
}
__device__ __forceinline__ const i4v* i4(const rt_node4* p) {
    return reinterpret_cast<const i4v*>(p);
}
__device__ __forceinline__ const RT_GLOBAL i4v* i4(const RT_GLOBAL rt_node4* p) {
    return reinterpret_cast<const RT_GLOBAL i4v*>(p);
}
template <typename T>
__device__ __forceinline__ RT_GLOBAL T* glb(T* p) {
    return (RT_GLOBAL T*)p;
}

// Keeps a node's child-index vector live before the "any child hit" branch:
// the compiler otherwise sinks that load into the branch, so a visit with a
// hit child pays a second memory round trip after the slab tests.  With the
// use here the load is issued with the other seven of the node.
__device__ __forceinline__ void keep_loaded(const i4v& v) {
    asm volatile("" : : "v"(v.x), "v"(v.y), "v"(v.z), "v"(v.w));
}

// The stack pointer is a separate local (a register); only the overflow
// array lives in private memory (keeping both in one aggregate would put the
// pointer in scratch too).
// The LDS column (`col`, entry e at col[e * bs], bs = block size) is derived
// from the __shared__ array itself (address space 3), never through a
// generic pointer: a select between an LDS and a private address would
// compile to a slower flat load.
template <int PRIV>
struct Stack {
    int32_t priv[PRIV];
};

// sp counts the entries; spo = sp * bs is the LDS slot offset of entry sp,
// kept by additions (a multiply per push / pop is a quarter-rate VALU op);
// lso = (ls - 1) * bs is the last LDS slot's.  The large private size serves
// trees of unknown depth (sized for kMaxDepth levels): there a push beyond
// the stack's capacity is dropped rather than written past it (only a tree
// deeper than kMaxDepth, which the entry points reject given its depth, can
// reach it).
#define RT_SPUSH(v)                                                      \
    do {                                                                 \
        if (PRIV == kPrivLarge && sp >= ls + PRIV) break;                \
        if (sp < ls) col[spo] = (v);                                     \
        else st.priv[sp - ls] = (v);                                     \
        ++sp;                                                            \
        spo += bs;                                                       \
    } while (0)

// Pop: the LDS read uses a clamped, always-valid slot (a ds_read); private
// memory is read only when the entry lives there.  The private read is
// volatile so it cannot be merged with the LDS read into one flat load of a
// selected address (which would put a flat-load latency on every pop).
#define RT_SPOP(dst)                                                     \
    do {                                                                 \
        --sp;                                                            \
        spo -= bs;                                                       \
        dst = col[min(spo, lso)];                                        \
        if (sp >= ls)                                                    \
            dst = reinterpret_cast<volatile int32_t*>(st.priv)[sp - ls]; \
    } while (0)

// The leaf children of one node, in their sorted order, as ONE loop of
// single-triangle steps: pend holds the sorted positions p (bit p) whose
// child is a leaf range.  Per ray this is the reference's sequence (each
// leaf child's range in sorted order, triangles in record order, the count
// added when a leaf child is started, any hit stopping at the first accepted
// triangle); per wave the loop runs max over lanes of the lane's triangle
// count instead of, position by position, the max over lanes of that
// position's count (divergent lanes share the one loop body).
// REC: hitPrim receives the accepted triangle's leaf record index instead of
// its triangle_id (for callers that use the hit only as an occlusion bit, or
// want the record).
template <bool ANY, bool COUNT, typename TP, bool REC = false>
__device__ __forceinline__ bool test_leaves(TP __restrict__ tris, const Child (&e)[4],
                                            uint32_t pend, float ox, float oy, float oz,
                                            float dx, float dy, float dz, float tmin,
                                            float& hitT, uint32_t& hitPrim,
                                            uint32_t& triTests) {
    uint32_t first = 0, left = 0;
    for (;;) {
        if (left == 0) {
            if (pend == 0) return false;
            const uint32_t k = __builtin_ctz(pend);
            pend &= pend - 1u;
            const int32_t c = k == 0 ? e[0].child : k == 1 ? e[1].child
                            : k == 2 ? e[2].child : e[3].child;
            const int32_t n = k == 0 ? e[0].count : k == 1 ? e[1].count
                            : k == 2 ? e[2].count : e[3].count;
            first = (uint32_t)~c;
            left = (uint32_t)n;
            if (COUNT) triTests += left;
        }
        // v0, v1 and v2.xy in two 16-B loads, v2.z (the record's ninth float)
        // in a 4-B one: the vector memory pipe is the traversal's bound
        const auto tp = f4(tris + first);
        const f4v a = tp[0], b = tp[1];
        const float c = f1(tris + first, 8);
        float t;
        if (mt_hit(a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c, ox, oy, oz, dx, dy, dz, &t) &&
            t >= tmin && t < hitT) {
            hitT = t;
            hitPrim = REC ? first : tris[first].triangle_id;
            if (ANY) return true;
        }
        ++first;
        --left;
    }
}

// A RADV Box32 node (include/radv_bvh.hpp:103-116: 4 child ids, then 4
// AABBs of min xyz, max xyz) as the four slots of a Box4 node.  A child id
// (offset >> 3) | type, offset from the BVH base (radv_bvh.hpp:39-49), of
// type box32 is an inner slot (count -1, child = offset / 128); of type
// triangle a one-triangle leaf slot (count 1, child = ~j for the triangle node
// j of the leaf section, which starts 128 boxes bytes after the base).  Empty
// ids (0xFFFFFFFF), other types, offsets outside their section and box
// children not after their parent (the root-first layout of every tree this
// library and fast_bvh_builder write) are empty slots, so a malformed image
// cannot send a load out of the image or form a cycle.
template <typename NP>
__device__ __forceinline__ i4v radv_node(NP p, int32_t self, uint32_t boxes, uint32_t ntris,
                                         float (&mnx)[4],
                                         float (&mxx)[4], float (&mny)[4], float (&mxy)[4],
                                         float (&mnz)[4], float (&mxz)[4], int32_t (&chv)[4],
                                         int32_t (&cnv)[4]) {
    const auto q = f4(p);
    const i4v ids = i4(p)[0];
    const f4v q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4], q5 = q[5], q6 = q[6];
    const float f[24] = {q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w,
                         q3.x, q3.y, q3.z, q3.w, q4.x, q4.y, q4.z, q4.w,
                         q5.x, q5.y, q5.z, q5.w, q6.x, q6.y, q6.z, q6.w};
    const int32_t idv[4] = {ids.x, ids.y, ids.z, ids.w};
    const uint64_t tbase = (uint64_t)boxes << 7;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        mnx[k] = f[6 * k];
        mny[k] = f[6 * k + 1];
        mnz[k] = f[6 * k + 2];
        mxx[k] = f[6 * k + 3];
        mxy[k] = f[6 * k + 4];
        mxz[k] = f[6 * k + 5];
        const uint32_t id = (uint32_t)idv[k], type = id & 7u;
        const uint64_t off = (uint64_t)(id & ~7u) << 3;
        const bool box = type == 5u && (off & 127u) == 0 && (off >> 7) < boxes &&
                         (off >> 7) > (uint64_t)self;
        const bool tri = type == 0u && off >= tbase && ((off - tbase) >> 6) < ntris &&
                         (off & 63u) == 0;
        chv[k] = box ? (int32_t)(off >> 7) : tri ? (int32_t)~(uint32_t)((off - tbase) >> 6) : 0;
        cnv[k] = box ? -1 : tri ? 1 : 0;
    }
    return ids;
}

// Loads rt_node4 ni and tests its four slots (software_bvh_bvh4.comp:116-144
// slab test + filter): entry t in te, hit slots in hit, child words and
// counts in chv / cnv; returns the child-index vector (kept live by the
// caller, see keep_loaded).  n?o / f?o: byte offsets of each axis's near and
// far planes in the node (min plane near for a positive inverse direction).
template <typename NP>
__device__ __forceinline__ i4v slab_f32(NP __restrict__ nodes, int32_t ni, uint32_t nxo,
                                        uint32_t nyo, uint32_t nzo, float ox, float oy,
                                        float oz, float ix, float iy, float iz, float tminc,
                                        float hitT, float (&te)[4], bool (&hit)[4],
                                        int32_t (&chv)[4], int32_t (&cnv)[4]) {
    const uint32_t fxo = nxo ^ 16u, fyo = nyo ^ 16u, fzo = nzo ^ 16u;
    const auto nb = bytes(nodes + ni);
    const f4v nx = *f4p(nb + nxo), fx = *f4p(nb + fxo), ny = *f4p(nb + nyo);
    const f4v fy = *f4p(nb + fyo), nz = *f4p(nb + nzo), fz = *f4p(nb + fzo);
    const i4v ch = i4(nodes + ni)[6];
    const i4v cn = i4(nodes + ni)[7];
    const float bnx[4] = {nx.x, nx.y, nx.z, nx.w}, bfx[4] = {fx.x, fx.y, fx.z, fx.w};
    const float bny[4] = {ny.x, ny.y, ny.z, ny.w}, bfy[4] = {fy.x, fy.y, fy.z, fy.w};
    const float bnz[4] = {nz.x, nz.y, nz.z, nz.w}, bfz[4] = {fz.x, fz.y, fz.z, fz.w};
    chv[0] = ch.x; chv[1] = ch.y; chv[2] = ch.z; chv[3] = ch.w;
    cnv[0] = cn.x; cnv[1] = cn.y; cnv[2] = cn.z; cnv[3] = cn.w;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float tnx = (bnx[k] - ox) * ix, tfx = (bfx[k] - ox) * ix;
        const float tny = (bny[k] - oy) * iy, tfy = (bfy[k] - oy) * iy;
        const float tnz = (bnz[k] - oz) * iz, tfz = (bfz[k] - oz) * iz;
        te[k] = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tminc));
        const float tx = fminf(fminf(tfx, tfy), fminf(tfz, hitT));
        hit[k] = cnv[k] != 0 && tx >= te[k];
    }
    return ch;
}

// After a node's slab tests, occlusion bits only (kOrderInner): inner hit
// children pushed nearest on top, leaf hit children tested in slot order;
// true when an accepted triangle ends the ray (hitPrim = its leaf record).  Rank r_k = inner hit
// children before k (larger entry t, ties by slot), so child k goes to stack
// entry sp + r_k: the near-first order of the inner children without
// sorting.  col / st / sp / spo / ls / bs: the ray's stack as in RT_SPUSH.
template <int PRIV, typename TP>
__device__ __forceinline__ bool inner_push_leaves(
    TP __restrict__ tris, const float (&te)[4], const bool (&hit)[4], const int32_t (&chv)[4],
    const int32_t (&cnv)[4], const i4v& ch, int32_t* col, Stack<PRIV>& st, int& sp, int& spo,
    int ls, int bs, float ox, float oy, float oz, float dx, float dy, float dz, float tmin,
    float& hitT, uint32_t& hitPrim) {
    bool inner[4];
    uint32_t leafm = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        inner[k] = hit[k] && cnv[k] < 0;
        leafm |= (hit[k] && cnv[k] > 0 ? 1u : 0u) << k;
    }
    keep_loaded(ch);
    if (!(hit[0] || hit[1] || hit[2] || hit[3])) return false;
    int rank[4] = {0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int k = j + 1; k < 4; ++k) {
            const bool jfirst = te[j] >= te[k];
            rank[k] += (inner[j] && jfirst) ? 1 : 0;
            rank[j] += (inner[k] && !jfirst) ? 1 : 0;
        }
    const int ni = (inner[0] ? 1 : 0) + (inner[1] ? 1 : 0) + (inner[2] ? 1 : 0) +
                   (inner[3] ? 1 : 0);
    if (sp + ni <= ls) {  // all four entries land in the LDS part
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (inner[k]) col[spo + rank[k] * bs] = chv[k];
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int q = sp + rank[k];
            // (large private part: sized for kMaxDepth levels; see RT_SPUSH)
            if (!inner[k] || (PRIV == kPrivLarge && q >= ls + PRIV)) continue;
            if (q < ls) col[q * bs] = chv[k];
            else st.priv[q - ls] = chv[k];
        }
    }
    const int nst = PRIV == kPrivLarge ? min(ni, ls + PRIV - sp) : ni;
    sp += nst;
    spo += nst * bs;
    if (!leafm) return false;
    const Child e[4] = {{te[0], chv[0], cnv[0]}, {te[1], chv[1], cnv[1]},
                        {te[2], chv[2], cnv[2]}, {te[3], chv[3], cnv[3]}};
    uint32_t triTests = 0;
    return test_leaves<true, false, TP, true>(tris, e, leafm, ox, oy, oz, dx, dy, dz, tmin, hitT,
                                              hitPrim, triTests);
}

// One ray's traversal (the body shared by k_trace, k_trace_batch,
// k_trace_scene and k_trace_radv).  col is the thread's LDS stack column (ls
// entries, stride bs).  Returns with hitT / hitPrim (closest: nearest
// accepted hit; any: first accepted hit).  RADV: nodes are the Box32 nodes of
// a RADV image (rboxes of them) and tris its triangle nodes (rtris); as the
// image is caller data, a ray also stops after rboxes pops (a tree pops each
// node at most once) and pushes beyond the stack's capacity are dropped.
// ORDER (any hit without counters or hit ids only): kOrderNear is the
// reference's order; kOrderFar pops the farthest hit child first; kOrderSlot
// skips the sort.  The occlusion bit of a ray does not depend on the order:
// until an accepted hit ends the ray, every node whose box meets [tmin, tmax]
// is visited, whatever the order.
template <int PRIV, bool ANY, bool COUNT, typename NP, typename TP, bool RADV = false,
          int ORDER = kOrderNear, bool Q8 = false>
__device__ __forceinline__ void trace_ray(NP __restrict__ nodes, TP __restrict__ tris, bool start,
                                          float ox, float oy, float oz, float tmin, float dx,
                                          float dy, float dz, int32_t* col, int ls, int bs,
                                          float& hitT, uint32_t& hitPrim, uint32_t& boxTests,
                                          uint32_t& triTests, uint32_t rboxes = 0,
                                          uint32_t rtris = 0) {
    // software_bvh_bvh4.comp:96-100
    const float ix = fabsf(dx) > 1e-6f ? 1.0f / dx : 1e6f;
    const float iy = fabsf(dy) > 1e-6f ? 1.0f / dy : 1e6f;
    const float iz = fabsf(dz) > 1e-6f ? 1.0f / dz : 1e6f;
    // keys of hit children are at least -FLT_MAX (below: missed children are
    // keyed -inf); clamping tmin once gives every key that bound, as
    // fmaxf(te, -FLT_MAX) per child would (max is exact)
    const float tminc = fmaxf(tmin, -FLT_MAX);
    // Byte offsets in a node of each axis's near and far planes: the min
    // plane is the near one for a positive inverse direction (ix is never 0).
    // As b_min <= b_max, (b_min - o) * ix <= (b_max - o) * ix for ix > 0 and
    // >= for ix < 0 under round-to-nearest, so the near plane's t is the
    // min of the two (and the far plane's the max) the reference takes: the
    // same values with no min / max per axis (a zero may differ in sign,
    // which no comparison sees).
    const uint32_t nxo = ix > 0.0f ? 0u : 16u, nyo = iy > 0.0f ? 32u : 48u;
    const uint32_t nzo = iz > 0.0f ? 64u : 80u;
    Stack<PRIV> st;
    int sp = 0, spo = 0;
    const int lso = (ls - 1) * bs;
    if (start) RT_SPUSH(0);
    uint32_t pops = 0;

    while (sp > 0) {
        if (RADV && ++pops > rboxes) break;
        int32_t ni;
        RT_SPOP(ni);
        int32_t chv[4], cnv[4];
        i4v ch;
        float te[4];
        bool hit[4];
        if constexpr (RADV) {
            float bminx[4], bmaxx[4], bminy[4], bmaxy[4], bminz[4], bmaxz[4];
            ch = radv_node(nodes + ni, ni, rboxes, rtris, bminx, bmaxx, bminy, bmaxy, bminz,
                           bmaxz, chv, cnv);
#pragma unroll
            for (int k = 0; k < 4; ++k) {  // :116-144 slab test + filter
                float t0x = (bminx[k] - ox) * ix, t1x = (bmaxx[k] - ox) * ix;
                float t0y = (bminy[k] - oy) * iy, t1y = (bmaxy[k] - oy) * iy;
                float t0z = (bminz[k] - oz) * iz, t1z = (bmaxz[k] - oz) * iz;
                float tminX = fminf(t0x, t1x), tmaxX = fmaxf(t0x, t1x);
                float tminY = fminf(t0y, t1y), tmaxY = fmaxf(t0y, t1y);
                float tminZ = fminf(t0z, t1z), tmaxZ = fmaxf(t0z, t1z);
                te[k] = fmaxf(fmaxf(tminX, tminY), fmaxf(tminZ, tminc));
                float tx = fminf(fminf(tmaxX, tmaxY), fminf(tmaxZ, hitT));
                hit[k] = cnv[k] != 0 && tx >= te[k];
            }
        } else if constexpr (Q8) {
            // rt_qnode4 (64 B, four 16-B loads): child boxes decoded on the
            // node's grid, p + q * 2^e (the product is exact, so the fma
            // rounds once, as the encoder's p + q * s does); the near / far
            // codes of an axis are its qlo / qhi words as for rt_node4's planes
            const auto qb = bytes(nodes) + (size_t)ni * 64u;
            const i4v q0 = i4p(qb)[0], q1 = i4p(qb)[1], q2 = i4p(qb)[2];
            ch = i4p(qb)[3];
            const float px = __int_as_float(q0.x), py = __int_as_float(q0.y);
            const float pz = __int_as_float(q0.z);
            const uint32_t ew = (uint32_t)q0.w;
            const float sx = __int_as_float((((int)(int8_t)(ew & 0xffu)) + 127) << 23);
            const float sy = __int_as_float((((int)(int8_t)((ew >> 8) & 0xffu)) + 127) << 23);
            const float sz = __int_as_float((((int)(int8_t)((ew >> 16) & 0xffu)) + 127) << 23);
            const uint32_t nqx = nxo ? (uint32_t)q1.w : (uint32_t)q1.x;
            const uint32_t fqx = nxo ? (uint32_t)q1.x : (uint32_t)q1.w;
            const uint32_t nqy = nyo == 32u ? (uint32_t)q1.y : (uint32_t)q2.x;
            const uint32_t fqy = nyo == 32u ? (uint32_t)q2.x : (uint32_t)q1.y;
            const uint32_t nqz = nzo == 64u ? (uint32_t)q1.z : (uint32_t)q2.y;
            const uint32_t fqz = nzo == 64u ? (uint32_t)q2.y : (uint32_t)q1.z;
            chv[0] = ch.x; chv[1] = ch.y; chv[2] = ch.z; chv[3] = ch.w;
            cnv[0] = (int32_t)(int16_t)(q2.z & 0xffff);
            cnv[1] = (int32_t)q2.z >> 16;
            cnv[2] = (int32_t)(int16_t)(q2.w & 0xffff);
            cnv[3] = (int32_t)q2.w >> 16;
#pragma unroll
            for (int k = 0; k < 4; ++k) {  // :116-144 slab test + filter, decoded boxes
                const int sh = 8 * k;
                const float bnx = __builtin_fmaf((float)((nqx >> sh) & 0xffu), sx, px);
                const float bfx = __builtin_fmaf((float)((fqx >> sh) & 0xffu), sx, px);
                const float bny = __builtin_fmaf((float)((nqy >> sh) & 0xffu), sy, py);
                const float bfy = __builtin_fmaf((float)((fqy >> sh) & 0xffu), sy, py);
                const float bnz = __builtin_fmaf((float)((nqz >> sh) & 0xffu), sz, pz);
                const float bfz = __builtin_fmaf((float)((fqz >> sh) & 0xffu), sz, pz);
                const float tnx = (bnx - ox) * ix, tfx = (bfx - ox) * ix;
                const float tny = (bny - oy) * iy, tfy = (bfy - oy) * iy;
                const float tnz = (bnz - oz) * iz, tfz = (bfz - oz) * iz;
                te[k] = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tminc));
                const float tx = fminf(fminf(tfx, tfy), fminf(tfz, hitT));
                hit[k] = cnv[k] != 0 && tx >= te[k];
            }
        } else {
            ch = slab_f32(nodes, ni, nxo, nyo, nzo, ox, oy, oz, ix, iy, iz, tminc, hitT, te, hit,
                          chv, cnv);
        }
        if (COUNT) boxTests += 4;

        if constexpr (ORDER == kOrderInner) {
            static_assert(ANY && !COUNT && !RADV, "for occlusion bits only");
            if (inner_push_leaves<PRIV>(tris, te, hit, chv, cnv, ch, col, st, sp, spo, ls, bs, ox,
                                        oy, oz, dx, dy, dz, tmin, hitT, hitPrim))
                break;  // the first accepted triangle ends the ray
            continue;
        }
        // :149-156 exchange sort of the hit children, run in place over the 4
        // slots with missed children keyed -inf (keys of hit children are at
        // least -FLT_MAX): for every pattern of hits and ties this yields the
        // reference's order of its compacted list, without compacting.
        Child e[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            e[k] = hit[k] ? Child{te[k], chv[k], cnv[k]} : Child{-INFINITY, 0, 0};
        keep_loaded(ch);
        if (!(hit[0] || hit[1] || hit[2] || hit[3])) continue;
        if constexpr (ORDER == kOrderNear) {
            cx(e[0], e[1]); cx(e[0], e[2]); cx(e[0], e[3]);
            cx(e[1], e[2]); cx(e[1], e[3]);
            cx(e[2], e[3]);
        } else if constexpr (ORDER == kOrderFar) {
            static_assert(ANY && !COUNT, "another order is for occlusion bits only");
            cx_up(e[0], e[1]); cx_up(e[0], e[2]); cx_up(e[0], e[3]);
            cx_up(e[1], e[2]); cx_up(e[1], e[3]);
            cx_up(e[2], e[3]);
        } else {
            static_assert(ANY && !COUNT, "another order is for occlusion bits only");
        }
        // :158-189 in sorted order: inner children pushed, leaf children
        // tested (padding, count 0, skipped).  The pushes do not depend on
        // the leaf tests, so all are made first and the leaves then run as
        // one loop (for any hit a push after the accepted leaf is never
        // popped: the ray ends).
        uint32_t pend = 0;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            if (e[p].count < 0 && (!RADV || sp < ls + PRIV)) RT_SPUSH(e[p].child);
            pend |= (e[p].count > 0 ? 1u : 0u) << p;
        }
        if (pend && test_leaves<ANY, COUNT>(tris, e, pend, ox, oy, oz, dx, dy, dz, tmin, hitT,
                                            hitPrim, triTests))
            break;  // any hit: the first accepted triangle ends the ray
    }
}

// Any-hit bits of a wave's 64 rays r0..r0+63 (r0 a multiple of 64): lane 0
// writes the first word, lane 32 the second, each if it holds a ray < n.
template <typename BP>
__device__ __forceinline__ void write_wave_bits(BP bits_out, uint32_t r0, uint32_t n,
                                                bool occluded) {
    const uint64_t m = __ballot(occluded);
    const uint32_t lane = threadIdx.x & 63u;
    if (lane == 0 && r0 < n) bits_out[r0 >> 5] = (uint32_t)m;
    if (lane == 32 && r0 + 32 < n) bits_out[(r0 >> 5) + 1] = (uint32_t)(m >> 32);
}

// Wave tiles of a ray grid (rt_trace_opts.grid_width / tile_width): the rays
// are rows of `width` consecutive rays (rows = n / width); wave w takes the
// tile of 2^tw_log2 rays across and 64 / 2^tw_log2 rows down, tiles in
// row-major order, so a wave's rays are a compact patch of the image instead
// of 64 rays of one row.
struct TileMap {
    uint32_t width;    // rays per row, a multiple of the tile width
    uint32_t rows;     // n / width
    uint32_t tw_log2;  // tile width 8, 16 or 32 as log2
};

// Ray of global thread g in the tile order; valid = inside the grid.
__device__ __forceinline__ uint32_t tile_ray(uint32_t g, const TileMap& T, bool& valid) {
    const uint32_t tile = g >> 6, lane = g & 63u;
    const uint32_t per_row = T.width >> T.tw_log2;
    const uint32_t trow = tile / per_row, tcol = tile - trow * per_row;
    const uint32_t y = (trow << (6u - T.tw_log2)) + (lane >> T.tw_log2);
    const uint32_t x = (tcol << T.tw_log2) + (lane & ((1u << T.tw_log2) - 1u));
    valid = y < T.rows;
    return valid ? y * T.width + x : 0u;
}

// Occlusion bits of a tiled wave: lanes 8k..8k+7 hold 8 consecutive rays of
// one row starting at a multiple of 8 (tile widths and the row width are
// multiples of 8), i.e. one byte of the bit array, stored by lane 8k.  The
// lane storing the last byte of the grid also clears the bytes after it up to
// the end of the last 32-bit word (as the linear kernel leaves them zero).
template <typename BP>
__device__ __forceinline__ void write_tile_bits(BP bits_out, uint32_t r, uint32_t n, bool valid,
                                                bool occluded) {
    const uint64_t m = __ballot(valid && occluded);
    const uint32_t lane = threadIdx.x & 63u;
    if (valid && (lane & 7u) == 0) {
        auto bytes = reinterpret_cast<uint8_t*>(bits_out);
        bytes[r >> 3] = (uint8_t)(m >> lane);
        if (r + 8 == n)
            for (uint32_t b = n >> 3; b < ((n + 31) >> 5) * 4; ++b) bytes[b] = 0;
    }
}

template <int PRIV, bool ANY, bool COUNT, int ORDER = kOrderNear, bool TILE = false,
          bool Q8 = false>
__global__ __launch_bounds__(kBlock) void k_trace(const rt_node4* __restrict__ nodes,
                                                  const rt_tri64* __restrict__ tris,
                                                  uint32_t node_count, rt_rays_soa rays,
                                                  uint32_t n, uint32_t ls,
                                                  uint32_t* __restrict__ prim_out,
                                                  float* __restrict__ t_out,
                                                  uint32_t* __restrict__ bits_out,
                                                  uint32_t* __restrict__ box_out,
                                                  uint32_t* __restrict__ tri_out,
                                                  TileMap tmap) {
    extern __shared__ int32_t lds_stack[];  // dynamic: ls * blockDim.x entries
    const uint32_t tid = threadIdx.x, bs = blockDim.x;
    bool valid;
    uint32_t r;
    if constexpr (TILE) {
        r = tile_ray(blockIdx.x * bs + tid, tmap, valid);
    } else {
        r = blockIdx.x * bs + tid;
        valid = r < n;
    }

    float ox = 0.0f, oy = 0.0f, oz = 0.0f, tmin = 1.0f;
    float dx = 0.0f, dy = 0.0f, dz = 1.0f, tmax = 0.0f;
    if (valid) {
        ox = rays.ox[r]; oy = rays.oy[r]; oz = rays.oz[r]; tmin = rays.tmin[r];
        dx = rays.dx[r]; dy = rays.dy[r]; dz = rays.dz[r]; tmax = rays.tmax[r];
    }
    float hitT = tmax;
    uint32_t hitPrim = RT_MISS;
    uint32_t boxTests = 0, triTests = 0;
    trace_ray<PRIV, ANY, COUNT, const rt_node4*, const rt_tri64*, false, ORDER, Q8>(
        nodes, tris, valid && node_count > 0, ox, oy, oz, tmin, dx, dy, dz, lds_stack + tid,
        (int)ls, (int)bs, hitT, hitPrim, boxTests, triTests);
    if (ANY) {
        if constexpr (TILE)
            write_tile_bits(bits_out, r, n, valid, hitPrim != RT_MISS);
        else
            write_wave_bits(bits_out, r - (tid & 63u), n, valid && hitPrim != RT_MISS);
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

// Packet traversal (RT_SCHED_PACKET), occlusion bits only: a wave walks one
// node order for its 64 rays (Garanzha & Loop 2010 trace coherent packets
// breadth-first; here depth-first with a per-entry lane mask).  Each stack
// entry carries the lanes whose own slab tests reached it, so a ray tests
// exactly the triangles its single-ray traversal reaches (same predicates),
// and its bit equals the single-ray kernel's whatever the order.  Node and
// leaf addresses are wave-uniform: node records come in through scalar
// loads, and a visit costs one slab test per lane instead of a vector load
// per lane of possibly 64 different nodes (the single-ray kernel's bound on
// shadow rays is the vector memory pipe: its loads are issued for the whole
// wave while about a third of the lanes are active).  Inner children are
// pushed nearest last by the entry t of the first lane that reached each.
constexpr int kPacketStack = 3 * (int)kMaxDepth + 1;  // live entries of a depth-64 tree

// A float read at a wave-uniform address, as a wave-uniform value.
__device__ __forceinline__ float uniform_f(const float& x) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}

template <bool TILE>
__global__ __launch_bounds__(kBlock) void k_trace_packet(const rt_node4* __restrict__ nodes,
                                                         const rt_tri64* __restrict__ tris,
                                                         uint32_t node_count, rt_rays_soa rays,
                                                         uint32_t n,
                                                         uint32_t* __restrict__ bits_out,
                                                         TileMap tmap) {
    extern __shared__ uint4 pstack[];  // dynamic: kPacketStack entries per wave
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    uint4* S = pstack + (tid >> 6) * kPacketStack;
    bool valid;
    uint32_t r;
    if constexpr (TILE) {
        r = tile_ray(blockIdx.x * blockDim.x + tid, tmap, valid);
    } else {
        r = blockIdx.x * blockDim.x + tid;
        valid = r < n;
    }
    float ox = 0.0f, oy = 0.0f, oz = 0.0f, tmin = 1.0f;
    float dx = 0.0f, dy = 0.0f, dz = 1.0f, tmax = 0.0f;
    if (valid) {
        ox = rays.ox[r]; oy = rays.oy[r]; oz = rays.oz[r]; tmin = rays.tmin[r];
        dx = rays.dx[r]; dy = rays.dy[r]; dz = rays.dz[r]; tmax = rays.tmax[r];
    }
    // the single-ray kernel's setup (software_bvh_bvh4.comp:96-100)
    const float ix = fabsf(dx) > 1e-6f ? 1.0f / dx : 1e6f;
    const float iy = fabsf(dy) > 1e-6f ? 1.0f / dy : 1e6f;
    const float iz = fabsf(dz) > 1e-6f ? 1.0f / dz : 1e6f;
    const float tminc = fmaxf(tmin, -FLT_MAX);
    const bool pxs = ix > 0.0f, pys = iy > 0.0f, pzs = iz > 0.0f;
    uint64_t alive = __ballot(valid && node_count > 0);  // lanes not yet occluded
    bool occl = false;
    int sp = 0;
    if (alive) {
        if (lane == 0) S[0] = make_uint4(0u, (uint32_t)alive, (uint32_t)(alive >> 32), 0u);
        sp = 1;
    }
    for (uint32_t pops = 0; sp > 0 && alive && pops < node_count; ++pops) {
        --sp;
        const uint4 e = S[sp];
        const uint32_t node = __builtin_amdgcn_readfirstlane(e.x);
        const uint64_t m = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane(e.z) << 32 |
                            (uint32_t)__builtin_amdgcn_readfirstlane(e.y)) & alive;
        if (m == 0) continue;
        const bool in = (m >> lane) & 1u;
        const rt_node4* nd = nodes + node;
        float te[4];
        uint64_t hm[4];
        int32_t chv[4], cnv[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {  // :116-144 slab test + filter, per lane
            // the 24 bounds as wave-uniform values first (a per-lane choice
            // of the load address would make them vector loads)
            chv[k] = (int32_t)__builtin_amdgcn_readfirstlane(nd->child[k]);
            cnv[k] = (int32_t)__builtin_amdgcn_readfirstlane(nd->count[k]);
            const float x0 = uniform_f(nd->minX[k]), x1 = uniform_f(nd->maxX[k]);
            const float y0 = uniform_f(nd->minY[k]), y1 = uniform_f(nd->maxY[k]);
            const float z0 = uniform_f(nd->minZ[k]), z1 = uniform_f(nd->maxZ[k]);
            const float bnx = pxs ? x0 : x1, bfx = pxs ? x1 : x0;
            const float bny = pys ? y0 : y1, bfy = pys ? y1 : y0;
            const float bnz = pzs ? z0 : z1, bfz = pzs ? z1 : z0;
            const float tnx = (bnx - ox) * ix, tfx = (bfx - ox) * ix;
            const float tny = (bny - oy) * iy, tfy = (bfy - oy) * iy;
            const float tnz = (bnz - oz) * iz, tfz = (bfz - oz) * iz;
            te[k] = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tminc));
            const float tx = fminf(fminf(tfx, tfy), fminf(tfz, tmax));
            hm[k] = __ballot(in && cnv[k] != 0 && tx >= te[k]);
        }
        // leaf children: each triangle by the lanes that reached it and are
        // not yet occluded (test_leaves' predicate: t >= tmin and t < tmax)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (cnv[k] <= 0 || !(hm[k] & alive)) continue;
            const uint32_t first = (uint32_t)~chv[k];
            for (int32_t j = 0; j < cnv[k]; ++j) {
                const uint64_t tm = hm[k] & alive;
                if (!tm) break;
                // every lane computes (uniform control flow: the record is a
                // scalar load); only the lanes of tm take the result
                const rt_tri64* tp = tris + first + j;
                float t;
                const bool h =
                    mt_hit(uniform_f(tp->v[0]), uniform_f(tp->v[1]), uniform_f(tp->v[2]),
                           uniform_f(tp->v[3]), uniform_f(tp->v[4]), uniform_f(tp->v[5]),
                           uniform_f(tp->v[6]), uniform_f(tp->v[7]), uniform_f(tp->v[8]), ox,
                           oy, oz, dx, dy, dz, &t) &&
                    t >= tmin && t < tmax;
                occl = occl || (h && ((tm >> lane) & 1u));
                alive &= ~__ballot(occl);
            }
        }
        // inner children reached by a live lane, nearest last: rank = the
        // children before it (a larger key, ties by slot) from the entry t
        // of the first lane reaching each
        bool inner[4];
        float key[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            inner[k] = cnv[k] < 0 && (hm[k] & alive) != 0;
            const int rep = inner[k] ? __builtin_ctzll(hm[k] & alive) : 0;
            key[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(te[k]), rep));
        }
        int rank[4] = {0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int k = j + 1; k < 4; ++k) {
                const bool jfirst = key[j] >= key[k];
                rank[k] += (inner[j] && jfirst) ? 1 : 0;
                rank[j] += (inner[k] && !jfirst) ? 1 : 0;
            }
        int ni = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (!inner[k]) continue;
            ++ni;
            const int q = sp + rank[k];
            const uint64_t cm = hm[k] & alive;
            if (lane == 0 && q < kPacketStack)  // (deeper than 64 levels: dropped)
                S[q] = make_uint4((uint32_t)chv[k], (uint32_t)cm, (uint32_t)(cm >> 32), 0u);
        }
        sp = min(sp + ni, kPacketStack);
    }
    if constexpr (TILE)
        write_tile_bits(bits_out, r, n, valid, occl);
    else
        write_wave_bits(bits_out, r - (tid & 63u), n, valid && occl);
}

// k_trace over a RADV image: nodes = the Box32 nodes (image + bvh_offset),
// tris = the triangle nodes (image + leaf_nodes_offset).
template <int PRIV, bool ANY, bool COUNT>
__global__ __launch_bounds__(kBlock) void k_trace_radv(const rt_node4* __restrict__ nodes,
                                                       const rt_tri64* __restrict__ tris,
                                                       uint32_t boxes, uint32_t ntris,
                                                       rt_rays_soa rays, uint32_t n, uint32_t ls,
                                                       uint32_t* __restrict__ prim_out,
                                                       float* __restrict__ t_out,
                                                       uint32_t* __restrict__ bits_out,
                                                       uint32_t* __restrict__ box_out,
                                                       uint32_t* __restrict__ tri_out) {
    extern __shared__ int32_t lds_stack[];
    const uint32_t tid = threadIdx.x, bs = blockDim.x;
    const uint32_t r = blockIdx.x * bs + tid;
    const bool valid = r < n;
    float ox = 0.0f, oy = 0.0f, oz = 0.0f, tmin = 1.0f;
    float dx = 0.0f, dy = 0.0f, dz = 1.0f, tmax = 0.0f;
    if (valid) {
        ox = rays.ox[r]; oy = rays.oy[r]; oz = rays.oz[r]; tmin = rays.tmin[r];
        dx = rays.dx[r]; dy = rays.dy[r]; dz = rays.dz[r]; tmax = rays.tmax[r];
    }
    float hitT = tmax;
    uint32_t hitPrim = RT_MISS;
    uint32_t boxTests = 0, triTests = 0;
    trace_ray<PRIV, ANY, COUNT, const rt_node4*, const rt_tri64*, true>(
        nodes, tris, valid && boxes > 0, ox, oy, oz, tmin, dx, dy, dz, lds_stack + tid, (int)ls,
        (int)bs, hitT, hitPrim, boxTests, triTests, boxes, ntris);
    if (ANY) {
        write_wave_bits(bits_out, r - (tid & 63u), n, valid && hitPrim != RT_MISS);
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

// Many (tree, rays, mode) items in one launch.  Item i owns the 64-ray chunks
// [chunk[i], chunk[i+1]) of the launch, so each wave has one item: its tree,
// rays and mode are wave-uniform (found by a binary search over chunk[]).
template <int PRIV>
__global__ __launch_bounds__(kBlock) void k_trace_batch(const rt_trace_item* __restrict__ items,
                                                        const uint32_t* __restrict__ chunk,
                                                        uint32_t count, uint32_t chunks,
                                                        uint32_t ls) {
    extern __shared__ int32_t lds_stack[];  // dynamic: ls * blockDim.x entries
    const uint32_t tid = threadIdx.x, lane = tid & 63u, bs = blockDim.x;
    const uint32_t c = __builtin_amdgcn_readfirstlane(blockIdx.x * (bs / 64) + (tid >> 6));
    if (c >= chunks) return;
    uint32_t lo = 0, hi = count;  // last item with chunk[item] <= c
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (chunk[mid] <= c) lo = mid;
        else hi = mid;
    }
    const rt_trace_item& it = items[lo];
    const uint32_t n = it.count;
    const uint32_t r0 = (c - chunk[lo]) * 64u;
    const uint32_t r = r0 + lane;
    const bool valid = r < n;
    const rt_rays_soa& rays = it.rays;
    float ox = 0.0f, oy = 0.0f, oz = 0.0f, tmin = 1.0f;
    float dx = 0.0f, dy = 0.0f, dz = 1.0f, tmax = 0.0f;
    if (valid) {
        ox = glb(rays.ox)[r]; oy = glb(rays.oy)[r]; oz = glb(rays.oz)[r];
        tmin = glb(rays.tmin)[r];
        dx = glb(rays.dx)[r]; dy = glb(rays.dy)[r]; dz = glb(rays.dz)[r];
        tmax = glb(rays.tmax)[r];
    }
    float hitT = tmax;
    uint32_t hitPrim = RT_MISS;
    uint32_t boxTests = 0, triTests = 0;
    const bool start = valid && it.bvh.node_count > 0;
    const auto nodes = glb(it.bvh.nodes);
    const auto tris = glb(it.bvh.tris);
    if (it.flags & RT_ANY) {
        trace_ray<PRIV, true, false>(nodes, tris, start, ox, oy, oz, tmin, dx, dy, dz,
                                     lds_stack + tid, (int)ls, (int)bs, hitT, hitPrim, boxTests,
                                     triTests);
        write_wave_bits(glb(it.occlusion_bits), r0, n, valid && hitPrim != RT_MISS);
        if (valid && it.hits.prim) glb(it.hits.prim)[r] = hitPrim;
    } else {
        trace_ray<PRIV, false, false>(nodes, tris, start, ox, oy, oz, tmin, dx, dy, dz,
                                      lds_stack + tid, (int)ls, (int)bs, hitT, hitPrim, boxTests,
                                      triTests);
        if (valid) {
            glb(it.hits.prim)[r] = hitPrim;
            glb(it.hits.t)[r] = hitT;
        }
    }
}

// Two-level scene: every ray visits the instances in order; instance i's
// BLAS is traversed on the mapped ray with hitT = the best t so far, so a
// later instance must be strictly nearer (ties keep the lower instance), and
// an any-hit ray stops at the first occluding instance.  Instance records
// are wave-uniform (scalar loads); the map is applied as written, in the
// order ((m0 x + m1 y) + m2 z) + t.
template <int PRIV, bool ANY, bool TILE = false>
__global__ __launch_bounds__(kBlock) void k_trace_scene(const rt_instance* __restrict__ inst,
                                                        uint32_t count, rt_rays_soa rays,
                                                        uint32_t n, uint32_t* __restrict__ prim_out,
                                                        float* __restrict__ t_out,
                                                        uint32_t* __restrict__ inst_out,
                                                        uint32_t* __restrict__ bits_out,
                                                        uint32_t ls, TileMap tmap) {
    extern __shared__ int32_t lds_stack[];  // dynamic: ls * blockDim.x entries
    const uint32_t tid = threadIdx.x, bs = blockDim.x;
    bool valid;
    uint32_t r;
    if constexpr (TILE) {
        r = tile_ray(blockIdx.x * bs + tid, tmap, valid);
    } else {
        r = blockIdx.x * bs + tid;
        valid = r < n;
    }
    float ox = 0.0f, oy = 0.0f, oz = 0.0f, tmin = 1.0f;
    float dx = 0.0f, dy = 0.0f, dz = 1.0f, tmax = 0.0f;
    if (valid) {
        ox = rays.ox[r]; oy = rays.oy[r]; oz = rays.oz[r]; tmin = rays.tmin[r];
        dx = rays.dx[r]; dy = rays.dy[r]; dz = rays.dz[r]; tmax = rays.tmax[r];
    }
    float best = tmax;
    uint32_t bprim = RT_MISS, binst = RT_MISS;
    for (uint32_t i = 0; i < count; ++i) {
        const rt_instance& I = inst[i];
        float mox = ox, moy = oy, moz = oz, mdx = dx, mdy = dy, mdz = dz;
        if (I.kind == RT_INST_TRANSLATE) {
            const float* m = I.world_to_object;
            mox = ox + m[3];
            moy = oy + m[7];
            moz = oz + m[11];
        } else if (I.kind == RT_INST_GENERAL) {
            const float* m = I.world_to_object;
            mox = m[0] * ox + m[1] * oy + m[2] * oz + m[3];
            moy = m[4] * ox + m[5] * oy + m[6] * oz + m[7];
            moz = m[8] * ox + m[9] * oy + m[10] * oz + m[11];
            mdx = m[0] * dx + m[1] * dy + m[2] * dz;
            mdy = m[4] * dx + m[5] * dy + m[6] * dz;
            mdz = m[8] * dx + m[9] * dy + m[10] * dz;
        }
        float hitT = best;
        uint32_t hitPrim = RT_MISS, bt = 0, tt = 0;
        trace_ray<PRIV, ANY, false>(glb(I.blas.nodes), glb(I.blas.tris),
                                    valid && I.blas.node_count > 0, mox, moy, moz, tmin, mdx, mdy,
                                    mdz, lds_stack + tid, (int)ls, (int)bs, hitT, hitPrim, bt,
                                    tt);
        if (hitPrim != RT_MISS) {
            best = hitT;
            bprim = hitPrim;
            binst = i;
            if (ANY) break;
        }
    }
    if (ANY) {
        if constexpr (TILE)
            write_tile_bits(bits_out, r, n, valid, bprim != RT_MISS);
        else
            write_wave_bits(bits_out, r - (tid & 63u), n, valid && bprim != RT_MISS);
    } else if (valid) {
        prim_out[r] = bprim;
        t_out[r] = best;
        if (inst_out) inst_out[r] = binst;
    }
}

// Refill form ("while-while" with lane refill): each wave owns the
// contiguous rays [w * per_wave, (w + 1) * per_wave) and, whenever at least
// `refill` lanes have finished, hands the next rays of its range to the idle
// lanes (ballot + mbcnt compaction), so a wave stays full while its slowest
// rays are still traversing; the hardware dispatcher balances the chunks
// across CUs.  Any-hit bits collect in the wave's LDS words (LDS atomicOr) and
// are written once at the end.  Same per-ray arithmetic and order as k_trace.
constexpr uint32_t kMaxPerWave = 4096;  // rays per wave: 128 occlusion words in LDS

// One wave's refill loop over rays [wbeg, end) of one ray set (wbeg a
// multiple of 64): whenever at least `refill` lanes are idle (or all are),
// the idle lanes take the next rays of the range (ballot + mbcnt
// compaction); each lane advances its ray by one node per iteration.  Any-hit
// bits collect in wbits (the wave's LDS words, zeroed here) and are flushed
// at the end.  Pointers are used as global (address space 1) views.
// BITS (any hit, no counters, no hit ids): the node step is the bits-only
// one of k_trace_refill_bits (near / far planes by offset, inner children at
// their ranks, leaf children in slot order); the bits do not depend on it.
template <int PRIV, bool ANY, bool COUNT, bool BITS = false>
__device__ __forceinline__ void refill_wave(
    const rt_node4* __restrict__ nodes_, const rt_tri64* __restrict__ tris_, uint32_t node_count,
    const rt_rays_soa& rays_, uint32_t wbeg32, uint32_t end, uint32_t refill, int32_t* col,
    int ls, int bs, uint32_t* wbits, uint32_t* prim_out_, float* t_out_, uint32_t* bits_out_,
    uint32_t* box_out_, uint32_t* tri_out_) {
    const auto nodes = glb(nodes_);
    const auto tris = glb(tris_);
    const auto rox = glb(rays_.ox), roy = glb(rays_.oy), roz = glb(rays_.oz);
    const auto rtmin = glb(rays_.tmin), rdx = glb(rays_.dx), rdy = glb(rays_.dy);
    const auto rdz = glb(rays_.dz), rtmax = glb(rays_.tmax);
    const auto prim_out = glb(prim_out_);
    const auto t_out = glb(t_out_);
    const auto bits_out = glb(bits_out_);
    const auto box_out = glb(box_out_);
    const auto tri_out = glb(tri_out_);
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t cur = wbeg32;
    if (ANY)
        for (uint32_t i = lane; i < kMaxPerWave / 32; i += 64) wbits[i] = 0u;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64u - lane));

    bool active = false;
    uint32_t r = 0;
    float ox = 0.f, oy = 0.f, oz = 0.f, tmin = 1.f, dx = 0.f, dy = 0.f, dz = 1.f;
    float ix = 0.f, iy = 0.f, iz = 0.f, hitT = 0.f;
    uint32_t hitPrim = RT_MISS, boxTests = 0, triTests = 0;
    float tminc = -FLT_MAX;  // fmaxf(tmin, -FLT_MAX), as in trace_ray
    uint32_t nxo = 0u, nyo = 32u, nzo = 64u;  // BITS: near planes by offset
    static_assert(!BITS || (ANY && !COUNT), "bits-only step: any hit without counters");
    Stack<PRIV> st;
    int sp = 0, spo = 0;
    const int lso = (ls - 1) * bs;

    for (;;) {
        const uint64_t idle = __ballot(!active);
        if (idle == ~0ull && cur >= end) break;
        if (cur < end && ((uint32_t)__popcll(idle) >= refill || idle == ~0ull)) {
            if (!active) {
                const uint32_t rr = cur + __popcll(idle & lt);
                if (rr < end) {
                    r = rr;
                    ox = rox[r]; oy = roy[r]; oz = roz[r]; tmin = rtmin[r];
                    dx = rdx[r]; dy = rdy[r]; dz = rdz[r]; hitT = rtmax[r];
                    ix = fabsf(dx) > 1e-6f ? 1.0f / dx : 1e6f;
                    iy = fabsf(dy) > 1e-6f ? 1.0f / dy : 1e6f;
                    iz = fabsf(dz) > 1e-6f ? 1.0f / dz : 1e6f;
                    tminc = fmaxf(tmin, -FLT_MAX);
                    if (BITS) {
                        nxo = ix > 0.0f ? 0u : 16u;
                        nyo = iy > 0.0f ? 32u : 48u;
                        nzo = iz > 0.0f ? 64u : 80u;
                    }
                    hitPrim = RT_MISS;
                    boxTests = triTests = 0;
                    sp = 0;
                    spo = 0;
                    if (node_count > 0) RT_SPUSH(0);
                    active = true;
                }
            }
            cur = (uint32_t)min<uint64_t>((uint64_t)cur + __popcll(idle), end);
        }
        if (!active) continue;

        bool finished = sp == 0;
        if (BITS && !finished) {
            int32_t ni;
            RT_SPOP(ni);
            float te[4];
            bool hit[4];
            int32_t chv[4], cnv[4];
            const i4v ch = slab_f32(nodes, ni, nxo, nyo, nzo, ox, oy, oz, ix, iy, iz, tminc,
                                    hitT, te, hit, chv, cnv);
            finished = inner_push_leaves<PRIV>(tris, te, hit, chv, cnv, ch, col, st, sp, spo, ls,
                                               bs, ox, oy, oz, dx, dy, dz, tmin, hitT, hitPrim) ||
                       sp == 0;
        } else if (!finished) {
            int32_t ni;
            RT_SPOP(ni);
            const auto np = f4(nodes + ni);
            const f4v mnx = np[0], mxx = np[1], mny = np[2], mxy = np[3], mnz = np[4],
                      mxz = np[5];
            const i4v ch = i4(nodes + ni)[6];
            const i4v cn = i4(nodes + ni)[7];
            if (COUNT) boxTests += 4;
            const float bminx[4] = {mnx.x, mnx.y, mnx.z, mnx.w};
            const float bmaxx[4] = {mxx.x, mxx.y, mxx.z, mxx.w};
            const float bminy[4] = {mny.x, mny.y, mny.z, mny.w};
            const float bmaxy[4] = {mxy.x, mxy.y, mxy.z, mxy.w};
            const float bminz[4] = {mnz.x, mnz.y, mnz.z, mnz.w};
            const float bmaxz[4] = {mxz.x, mxz.y, mxz.z, mxz.w};
            const int32_t chv[4] = {ch.x, ch.y, ch.z, ch.w};
            const int32_t cnv[4] = {cn.x, cn.y, cn.z, cn.w};
            float te[4];
            bool hit[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                float t0x = (bminx[k] - ox) * ix, t1x = (bmaxx[k] - ox) * ix;
                float t0y = (bminy[k] - oy) * iy, t1y = (bmaxy[k] - oy) * iy;
                float t0z = (bminz[k] - oz) * iz, t1z = (bmaxz[k] - oz) * iz;
                float tminX = fminf(t0x, t1x), tmaxX = fmaxf(t0x, t1x);
                float tminY = fminf(t0y, t1y), tmaxY = fmaxf(t0y, t1y);
                float tminZ = fminf(t0z, t1z), tmaxZ = fmaxf(t0z, t1z);
                te[k] = fmaxf(fmaxf(tminX, tminY), fmaxf(tminZ, tminc));
                float tx = fminf(fminf(tmaxX, tmaxY), fminf(tmaxZ, hitT));
                hit[k] = cnv[k] != 0 && tx >= te[k];
            }
            Child e[4];  // in-place exchange sort, as in k_trace
#pragma unroll
            for (int k = 0; k < 4; ++k)
                e[k] = hit[k] ? Child{te[k], chv[k], cnv[k]} : Child{-INFINITY, 0, 0};
            keep_loaded(ch);
            if (hit[0] || hit[1] || hit[2] || hit[3]) {
                cx(e[0], e[1]); cx(e[0], e[2]); cx(e[0], e[3]);
                cx(e[1], e[2]); cx(e[1], e[3]);
                cx(e[2], e[3]);
                uint32_t pend = 0;  // as trace_ray: pushes, then the leaves as one loop
#pragma unroll
                for (int p = 0; p < 4; ++p) {
                    if (e[p].count < 0) RT_SPUSH(e[p].child);
                    pend |= (e[p].count > 0 ? 1u : 0u) << p;
                }
                if (pend && test_leaves<ANY, COUNT>(tris, e, pend, ox, oy, oz, dx, dy, dz, tmin,
                                                    hitT, hitPrim, triTests))
                    finished = true;
            }
            if (sp == 0) finished = true;
        }
        if (finished) {
            if (ANY) {
                if (hitPrim != RT_MISS) atomicOr(wbits + ((r - wbeg32) >> 5), 1u << (r & 31u));
                if (prim_out) prim_out[r] = hitPrim;
            } else {
                prim_out[r] = hitPrim;
                t_out[r] = hitT;
            }
            if (COUNT) {
                box_out[r] = boxTests;
                tri_out[r] = triTests;
            }
            active = false;
        }
    }
    if (ANY) {  // flush this wave's occlusion words (whole words: no overlap)
        const uint32_t words = (end - wbeg32 + 31) >> 5;
        for (uint32_t i = lane; i < words; i += 64) bits_out[(wbeg32 >> 5) + i] = wbits[i];
    }
}

template <int PRIV, bool ANY, bool COUNT>
__global__ __launch_bounds__(kBlock) void k_trace_persist(
    const rt_node4* __restrict__ nodes, const rt_tri64* __restrict__ tris, uint32_t node_count,
    rt_rays_soa rays, uint32_t n, uint32_t per_wave, uint32_t refill, uint32_t lsu,
    uint32_t* __restrict__ prim_out, float* __restrict__ t_out, uint32_t* __restrict__ bits_out,
    uint32_t* __restrict__ box_out, uint32_t* __restrict__ tri_out) {
    extern __shared__ int32_t lds_stack[];  // dynamic: ls * blockDim.x entries
    // any hit: each wave's occlusion words (its range is whole 32-ray words)
    __shared__ uint32_t lds_bits[ANY ? (kBlock / 64) * (kMaxPerWave / 32) : 1];
    const uint32_t tid = threadIdx.x;
    const uint32_t wave = (blockIdx.x * blockDim.x + tid) >> 6;
    const uint64_t wbeg = (uint64_t)wave * per_wave;
    const uint32_t beg = (uint32_t)min<uint64_t>(wbeg, n);  // a multiple of 64 (per_wave is)
    const uint32_t end = (uint32_t)min<uint64_t>(wbeg + per_wave, n);
    refill_wave<PRIV, ANY, COUNT>(nodes, tris, node_count, rays, beg, end, refill,
                                  lds_stack + tid, (int)lsu, (int)blockDim.x,
                                  lds_bits + (ANY ? (tid >> 6) * (kMaxPerWave / 32) : 0),
                                  prim_out, t_out, bits_out, box_out, tri_out);
}

// Refill form for occlusion bits only (no counters, no hit ids), with or
// without wave tiles: the node step of k_trace's bits-only path (near / far
// planes by offset, inner children at their ranks, leaf children in slot
// order) inside refill_wave's loop.  Wave w owns the launch's thread slots
// [w * per_wave, (w + 1) * per_wave) in k_trace's order (with tiles: tile
// by tile, each tile's 64 slots row by row; slots of a partial tile row
// below the grid hold no ray), so the rays a wave refills from are its next
// tiles.  Each finished ray's bit goes to the wave's LDS words at its slot;
// at the end every 8 slots (8 consecutive rays of one row: tile widths and
// the row width are multiples of 8) become one byte of the bit array, as in
// write_tile_bits.  A ray's bit does not depend on when it is traced.
template <int PRIV, bool TILE>
__global__ __launch_bounds__(kBlock) void k_trace_refill_bits(
    const rt_node4* __restrict__ nodes, const rt_tri64* __restrict__ tris, uint32_t node_count,
    rt_rays_soa rays, uint32_t n, uint32_t slots, uint32_t per_wave, uint32_t refill,
    uint32_t lsu, uint32_t* __restrict__ bits_out, TileMap tmap) {
    // dynamic: lsu * blockDim.x stack entries, then per_wave / 32 words per wave
    extern __shared__ int32_t lds_stack[];
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const int bs = (int)blockDim.x, ls = (int)lsu;
    int32_t* col = lds_stack + tid;
    uint32_t* wbits = reinterpret_cast<uint32_t*>(lds_stack + ls * bs) + (tid >> 6) * (per_wave >> 5);
    const uint64_t w0 = (uint64_t)((blockIdx.x * blockDim.x + tid) >> 6) * per_wave;
    const uint32_t beg = (uint32_t)min<uint64_t>(w0, slots);  // a multiple of 64
    const uint32_t end = (uint32_t)min<uint64_t>(w0 + per_wave, slots);
    for (uint32_t i = lane; i < (per_wave >> 5); i += 64) wbits[i] = 0u;

    bool active = false;
    uint32_t slot = 0;  // the lane's ray: slot - beg in the wave's words
    float ox = 0.f, oy = 0.f, oz = 0.f, tmin = 1.f, dx = 0.f, dy = 0.f, dz = 1.f;
    float ix = 0.f, iy = 0.f, iz = 0.f, hitT = 0.f, tminc = -FLT_MAX;
    uint32_t nxo = 0u, nyo = 32u, nzo = 64u, hitPrim = RT_MISS;
    Stack<PRIV> st;
    int sp = 0, spo = 0;
    const int lso = (ls - 1) * bs;
    uint32_t cur = beg;

    for (;;) {
        const uint64_t idle = __ballot(!active);
        if (idle == ~0ull && cur >= end) break;
        if (cur < end && ((uint32_t)__popcll(idle) >= refill || idle == ~0ull)) {
            if (!active) {
                // idle lanes below this one: its place among the new rays
                const uint32_t below = __builtin_amdgcn_mbcnt_hi(
                    (uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
                const uint32_t g = cur + below;
                bool valid = g < end;
                uint32_t r = g;
                if constexpr (TILE) {
                    if (valid) r = tile_ray(g, tmap, valid);
                }
                if (valid && node_count > 0) {
                    slot = g;
                    ox = rays.ox[r]; oy = rays.oy[r]; oz = rays.oz[r]; tmin = rays.tmin[r];
                    dx = rays.dx[r]; dy = rays.dy[r]; dz = rays.dz[r]; hitT = rays.tmax[r];
                    // software_bvh_bvh4.comp:96-100, planes as in trace_ray
                    ix = fabsf(dx) > 1e-6f ? 1.0f / dx : 1e6f;
                    iy = fabsf(dy) > 1e-6f ? 1.0f / dy : 1e6f;
                    iz = fabsf(dz) > 1e-6f ? 1.0f / dz : 1e6f;
                    tminc = fmaxf(tmin, -FLT_MAX);
                    nxo = ix > 0.0f ? 0u : 16u;
                    nyo = iy > 0.0f ? 32u : 48u;
                    nzo = iz > 0.0f ? 64u : 80u;
                    hitPrim = RT_MISS;
                    sp = 0;
                    spo = 0;
                    RT_SPUSH(0);
                    active = true;
                }
            }
            cur = (uint32_t)min<uint64_t>((uint64_t)cur + __popcll(idle), end);
        }
        if (!active) continue;

        bool ended = false;
        if (sp > 0) {
            int32_t ni;
            RT_SPOP(ni);
            float te[4];
            bool hit[4];
            int32_t chv[4], cnv[4];
            const i4v ch = slab_f32(nodes, ni, nxo, nyo, nzo, ox, oy, oz, ix, iy, iz, tminc,
                                    hitT, te, hit, chv, cnv);
            ended = inner_push_leaves<PRIV>(tris, te, hit, chv, cnv, ch, col, st, sp, spo, ls, bs,
                                            ox, oy, oz, dx, dy, dz, tmin, hitT, hitPrim);
        }
        if (ended || sp == 0) {
            if (ended) atomicOr(wbits + ((slot - beg) >> 5), 1u << ((slot - beg) & 31u));
            active = false;
        }
    }
    wave_lds_sync();  // the wave's atomics before its flush
    if constexpr (TILE) {
        auto bytes_out = reinterpret_cast<uint8_t*>(bits_out);
        for (uint32_t m = lane; m < ((end - beg) >> 3); m += 64) {
            bool valid;
            const uint32_t r0 = tile_ray(beg + 8u * m, tmap, valid);
            if (!valid) continue;
            bytes_out[r0 >> 3] = (uint8_t)(wbits[m >> 2] >> ((m & 3u) * 8u));
            // the grid's last byte: clear the bytes after it up to the end of
            // the last 32-bit word (as the linear kernels leave them zero)
            if (r0 + 8 == n)
                for (uint32_t b = n >> 3; b < ((n + 31) >> 5) * 4; ++b) bytes_out[b] = 0;
        }
    } else {  // whole words (beg is a multiple of 64; bits past n stay zero)
        for (uint32_t i = lane; i < ((end - beg + 31) >> 5); i += 64)
            bits_out[(beg >> 5) + i] = wbits[i];
    }
}

// Occluder hints (rt_trace_occluded_hinted): occlusion bits only.
//
// For a tree whose slot boxes contain the boxes of the triangles under them
// (every tree this library builds or refits, and a SoftwareBVH4 upload), the
// traversal's occlusion bit of a ray is 1 exactly when some triangle T passes
// the predicate (t in [tmin, tmax)) while the box of T's leaf slot passes the
// slab test.  The slab test is monotone in the box: under round-to-nearest a
// box containing another has a near t no later and a far t no earlier, so
// every slot box on the way down to T's slot passes as well; T is then
// reached unless another accepted triangle ends the ray first, and either way
// the bit is 1.  T's own box (the min / max of its vertices) lies in its slot
// box, so a triangle that passes the predicate and whose own box passes the
// slab test proves the bit without a traversal: the hint test.  A ray whose
// hints all fail is traced by the bits-only node step of k_trace_refill_bits,
// so its bit is the traversal's.
struct HintArgs {
    const uint32_t* prim;    // hint triangle id per slot, RT_MISS for none
    const uint32_t* record;  // triangle id -> leaf record, RT_MISS for none
    uint32_t period;         // ray r's hint: prim[r % period]
    uint32_t ids;            // entries of record
    uint32_t wave;           // also try the hints of the other rays of the 64
};

// The hint test: leaf record rec through the leaf predicate of test_leaves,
// then its own box through the slab test of slab_f32 (min / max of t0, t1
// per axis: the same values as its near / far plane selection).
template <typename TP>
__device__ __forceinline__ bool hint_occludes(TP __restrict__ tris, uint32_t rec, float ox,
                                              float oy, float oz, float dx, float dy, float dz,
                                              float ix, float iy, float iz, float tmin,
                                              float tminc, float tmax) {
    const auto tp = f4(tris + rec);
    const f4v a = tp[0], b = tp[1];
    const float c = f1(tris + rec, 8);
    float t;
    if (!mt_hit(a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c, ox, oy, oz, dx, dy, dz, &t) ||
        !(t >= tmin && t < tmax))
        return false;
    // vertices v0 = a.xyz, v1 = (a.w, b.x, b.y), v2 = (b.z, b.w, c)
    const float mnx = fminf(fminf(a.x, a.w), b.z), mxx = fmaxf(fmaxf(a.x, a.w), b.z);
    const float mny = fminf(fminf(a.y, b.x), b.w), mxy = fmaxf(fmaxf(a.y, b.x), b.w);
    const float mnz = fminf(fminf(a.z, b.y), c), mxz = fmaxf(fmaxf(a.z, b.y), c);
    const float t0x = (mnx - ox) * ix, t1x = (mxx - ox) * ix;
    const float t0y = (mny - oy) * iy, t1y = (mxy - oy) * iy;
    const float t0z = (mnz - oz) * iz, t1z = (mxz - oz) * iz;
    const float te = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tminc));
    const float tx = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), tmax));
    return tx >= te;
}

// One ray of the launch: slot g's ray (the slot itself, or its wave-tile ray).
struct HRay {
    float ox, oy, oz, tmin, dx, dy, dz, tmax, ix, iy, iz, tminc;
};
template <bool TILE>
__device__ __forceinline__ HRay load_hray(const rt_rays_soa& rays, uint32_t g,
                                          const TileMap& tmap) {
    bool valid = true;
    uint32_t r = g;
    if constexpr (TILE) r = tile_ray(g, tmap, valid);
    HRay q;
    q.ox = rays.ox[r]; q.oy = rays.oy[r]; q.oz = rays.oz[r]; q.tmin = rays.tmin[r];
    q.dx = rays.dx[r]; q.dy = rays.dy[r]; q.dz = rays.dz[r]; q.tmax = rays.tmax[r];
    // software_bvh_bvh4.comp:96-100
    q.ix = fabsf(q.dx) > 1e-6f ? 1.0f / q.dx : 1e6f;
    q.iy = fabsf(q.dy) > 1e-6f ? 1.0f / q.dy : 1e6f;
    q.iz = fabsf(q.dz) > 1e-6f ? 1.0f / q.dz : 1e6f;
    q.tminc = fmaxf(q.tmin, -FLT_MAX);
    return q;
}

// The passes of rt_trace_occluded_hinted share an open mask over the launch's
// thread slots (k_trace's order; with wave tiles, tile by tile): bit g set =
// slot g's ray is still open.  The slots are cut into segments of
// seg_words * 32 slots; the trace pass streams the open rays segment after
// segment.
struct HintPlan {
    uint32_t* open;      // nseg * seg_words words
    uint32_t* counter;   // kXcds segment counters of the trace pass
    uint32_t nseg;       // segments
    uint32_t seg_words;  // open-mask words per segment (even)
    uint32_t layer_segs; // segments per layer of `period` rays (0: one queue)
    uint32_t layers;     // nseg / layer_segs
};

// Compute dies of an MI355X: workgroups of a launch go to them round-robin
// (workgroup b to die b % kXcds), each die with its own L2.
constexpr uint32_t kXcds = 8;

// Segment k of die x's queue.  With layers (the rays are layers of `period`
// rays, e.g. a frame's shadow rays light by light, each layer an image in
// slot order) die x takes band x of every layer: the segments
// [ceil(x B / 8), ceil((x + 1) B / 8)) of each layer of B segments, layer by
// layer, so a die's L2 holds the geometry of one image band.  Without
// layers there is one queue (die 0's), in slot order.  RT_MISS: k is past
// the queue's end.
__device__ __forceinline__ uint32_t queue_segment(const HintPlan& pl, uint32_t x, uint32_t k) {
    if (pl.layer_segs == 0) return x == 0 && k < pl.nseg ? k : RT_MISS;
    const uint32_t B = pl.layer_segs;
    const uint32_t b0 = (x * B + kXcds - 1) / kXcds, b1 = ((x + 1) * B + kXcds - 1) / kXcds;
    const uint32_t c = b1 - b0;
    if (c == 0 || k >= c * pl.layers) return RT_MISS;
    const uint32_t layer = k / c;
    return layer * B + b0 + (k - layer * c);
}

// Pass 1, one wave per 64 slots: each ray tries its own hint and, with
// h.wave, the distinct hints of the other 63 rays (one triangle for the whole
// wave at a time).  Writes every bit of the wave's rays (1 where a hint
// settles the ray) and the open mask of its 64 slots; block 0 zeroes the
// trace pass's segment counter.
template <bool TILE>
__global__ __launch_bounds__(kBlock) void k_hint_pass(const rt_tri64* __restrict__ tris,
                                                      uint32_t node_count, uint32_t tri_count,
                                                      rt_rays_soa rays, uint32_t n, uint32_t slots,
                                                      uint32_t* __restrict__ bits_out,
                                                      TileMap tmap, HintArgs h, HintPlan pl) {
    const uint32_t lane = threadIdx.x & 63u;
    if (blockIdx.x == 0 && threadIdx.x < kXcds) pl.counter[threadIdx.x] = 0u;
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t g0 = g - lane;
    if (g0 >= slots) {  // the last segment's padding (whole waves): closed
        if (lane == 0 || lane == 32) pl.open[(g0 >> 5) + (lane >> 5)] = 0u;
        return;
    }
    bool valid;
    uint32_t r = g;
    if constexpr (TILE) r = tile_ray(g, tmap, valid);
    else valid = g < n;
    const bool live = valid && node_count > 0;
    HRay q{0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 1.f, 0.f, 1e6f, 1e6f, 1.f, 1.f};
    uint32_t rec = RT_MISS;
    if (live) {
        q = load_hray<TILE>(rays, g, tmap);
        const uint32_t pid = h.prim[r % h.period];
        if (pid < h.ids) {
            const uint32_t c = h.record[pid];
            if (c < tri_count) rec = c;
        }
    }
    bool occl = rec != RT_MISS && hint_occludes(tris, rec, q.ox, q.oy, q.oz, q.dx, q.dy, q.dz,
                                                q.ix, q.iy, q.iz, q.tmin, q.tminc, q.tmax);
    if (h.wave) {
        // the wave's distinct hints in lane order, each tried by the rays
        // still open (a ray's own hint is not tried twice)
        uint64_t cand = __ballot(rec != RT_MISS);
        while (cand != 0ull && __ballot(live && !occl) != 0ull) {
            const uint32_t c = __builtin_amdgcn_readlane(rec, (int)__builtin_ctzll(cand));
            cand &= ~__ballot(rec == c);
            if (live && !occl && rec != c)
                occl = hint_occludes(tris, c, q.ox, q.oy, q.oz, q.dx, q.dy, q.dz, q.ix, q.iy, q.iz,
                                     q.tmin, q.tminc, q.tmax);
        }
    }
    if constexpr (TILE) write_tile_bits(bits_out, r, n, valid, occl);
    else write_wave_bits(bits_out, g0, n, valid && occl);
    const uint64_t om = __ballot(live && !occl);
    if (lane == 0) pl.open[g0 >> 5] = (uint32_t)om;
    if (lane == 32) pl.open[(g0 >> 5) + 1] = (uint32_t)(om >> 32);
}

constexpr uint32_t kRing = 128;  // staged rays per wave (a power of two)

// The trace pass, persistent: each wave takes segment after segment (an
// atomic counter per die, queue_segment; a die whose queue is empty takes
// from the others) and stages the segment's open rays in slot order, 64 slots
// at a time, into its ring of kRing entries; whenever `refill` lanes are idle
// they take the next staged rays (the refill loop of k_trace_refill_bits), so
// a wave stays filled across segment boundaries until no segment is left.  A
// ray that ends on an accepted triangle ORs its bit into bits_out.
template <int PRIV, bool TILE>
__global__ __launch_bounds__(kBlock) void k_hint_trace(const rt_node4* __restrict__ nodes,
                                                       const rt_tri64* __restrict__ tris,
                                                       rt_rays_soa rays, TileMap tmap,
                                                       HintPlan pl, uint32_t refill, uint32_t lsu,
                                                       uint32_t* __restrict__ bits_out) {
    // dynamic: lsu * blockDim.x stack entries, then per wave kRing slots
    extern __shared__ int32_t lds_stack[];
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const int bs = (int)blockDim.x, ls = (int)lsu;
    int32_t* col = lds_stack + tid;
    uint32_t* ring = reinterpret_cast<uint32_t*>(lds_stack + ls * bs) + (tid >> 6) * kRing;
    uint32_t* const counter = pl.counter;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64u - lane));
    bool active = false;
    uint32_t r = 0;  // the lane's ray
    HRay q{0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 0.f, -FLT_MAX};
    float hitT = 0.f;
    uint32_t nxo = 0u, nyo = 32u, nzo = 64u, hitPrim = RT_MISS;
    Stack<PRIV> st;
    int sp = 0, spo = 0;
    const int lso = (ls - 1) * bs;
    // wave-uniform: current segment, next word in it, staged range
    // [head, tail), whether segments may remain
    uint32_t seg = 0, wcur = pl.seg_words, head = 0, tail = 0;
    bool more = true;
    for (;;) {
        while (tail - head < 64u && more) {
            if (wcur >= pl.seg_words) {
                // the next segment of this die's queue, then of the others'
                uint32_t sg = RT_MISS;
                for (uint32_t j = 0; j < kXcds && sg == RT_MISS; ++j) {
                    const uint32_t x = (blockIdx.x + j) % kXcds;
                    if (pl.layer_segs == 0 && x != 0) continue;
                    uint32_t k = 0;
                    if (lane == 0) k = atomicAdd(counter + x, 1u);
                    sg = queue_segment(pl, x, __builtin_amdgcn_readfirstlane(k));
                }
                if (sg == RT_MISS) {
                    more = false;
                    break;
                }
                seg = sg;
                wcur = 0;
            }
            const uint32_t w = seg * pl.seg_words + wcur;
            const uint64_t m = (uint64_t)pl.open[w] | (uint64_t)pl.open[w + 1] << 32;
            if ((m >> lane) & 1ull)
                ring[(tail + (uint32_t)__popcll(m & lt)) & (kRing - 1)] = w * 32u + lane;
            tail += (uint32_t)__popcll(m);
            wcur += 2;
        }
        const uint64_t idle = __ballot(!active);
        if (idle == ~0ull && head == tail && !more) break;
        if (head < tail && ((uint32_t)__popcll(idle) >= refill || idle == ~0ull || !more)) {
            if (!active) {
                const uint32_t below = __builtin_amdgcn_mbcnt_hi(
                    (uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
                const uint32_t k = head + below;
                if (k < tail) {
                    const uint32_t g = ring[k & (kRing - 1)];
                    bool valid = true;
                    r = g;
                    if constexpr (TILE) r = tile_ray(g, tmap, valid);
                    q = load_hray<TILE>(rays, g, tmap);
                    hitT = q.tmax;
                    nxo = q.ix > 0.0f ? 0u : 16u;
                    nyo = q.iy > 0.0f ? 32u : 48u;
                    nzo = q.iz > 0.0f ? 64u : 80u;
                    hitPrim = RT_MISS;
                    sp = 0;
                    spo = 0;
                    RT_SPUSH(0);
                    active = true;
                }
            }
            head = min(head + (uint32_t)__popcll(idle), tail);
        }
        if (!active) continue;

        bool ended = false;
        if (sp > 0) {
            int32_t ni;
            RT_SPOP(ni);
            float te[4];
            bool hit[4];
            int32_t chv[4], cnv[4];
            const i4v ch = slab_f32(nodes, ni, nxo, nyo, nzo, q.ox, q.oy, q.oz, q.ix, q.iy, q.iz,
                                    q.tminc, hitT, te, hit, chv, cnv);
            ended = inner_push_leaves<PRIV>(tris, te, hit, chv, cnv, ch, col, st, sp, spo, ls, bs,
                                            q.ox, q.oy, q.oz, q.dx, q.dy, q.dz, q.tmin, hitT,
                                            hitPrim);
        }
        if (ended || sp == 0) {
            if (ended) atomicOr(bits_out + (r >> 5), 1u << (r & 31u));
            active = false;
        }
    }
}

// record[id] = RT_MISS for every id, then record[triangle_id of leaf i] = i.
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t count, uint32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}
__global__ void k_prim_records(const rt_tri64* __restrict__ tris, uint32_t tri_count,
                               uint32_t ids, uint32_t* __restrict__ record) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= tri_count) return;
    const uint32_t id = tris[i].triangle_id;
    if (id < ids) record[id] = i;
}

// Batched trace, refill form: item i owns the chunks [chunk[i], chunk[i+1])
// of per_wave rays each (rt_trace_batch_plan_ex with chunk_rays = per_wave);
// each wave takes one chunk (tree, rays and mode wave-uniform) and walks it
// with the refill loop of k_trace_persist.
template <int PRIV>
__global__ __launch_bounds__(kBlock) void k_trace_batch_refill(
    const rt_trace_item* __restrict__ items, const uint32_t* __restrict__ chunk, uint32_t count,
    uint32_t chunks, uint32_t per_wave, uint32_t refill, uint32_t ls) {
    extern __shared__ int32_t lds_stack[];  // dynamic: ls * blockDim.x entries
    __shared__ uint32_t lds_bits[(kBlock / 64) * (kMaxPerWave / 32)];
    const uint32_t tid = threadIdx.x, bs = blockDim.x;
    const uint32_t c = __builtin_amdgcn_readfirstlane(blockIdx.x * (bs / 64) + (tid >> 6));
    if (c >= chunks) return;
    uint32_t lo = 0, hi = count;  // last item with chunk[item] <= c
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (chunk[mid] <= c) lo = mid;
        else hi = mid;
    }
    const rt_trace_item& it = items[lo];
    const uint64_t b0 = (uint64_t)(c - chunk[lo]) * per_wave;
    const uint32_t beg = (uint32_t)min<uint64_t>(b0, it.count);
    const uint32_t end = (uint32_t)min<uint64_t>(b0 + per_wave, it.count);
    uint32_t* wbits = lds_bits + (tid >> 6) * (kMaxPerWave / 32);
    if ((it.flags & RT_ANY) && !it.hits.prim)  // occlusion bits only
        refill_wave<PRIV, true, false, true>(it.bvh.nodes, it.bvh.tris, it.bvh.node_count,
                                             it.rays, beg, end, refill, lds_stack + tid, (int)ls,
                                             (int)bs, wbits, nullptr, nullptr, it.occlusion_bits,
                                             nullptr, nullptr);
    else if (it.flags & RT_ANY)
        refill_wave<PRIV, true, false>(it.bvh.nodes, it.bvh.tris, it.bvh.node_count, it.rays,
                                       beg, end, refill, lds_stack + tid, (int)ls, (int)bs,
                                       wbits, it.hits.prim, nullptr, it.occlusion_bits, nullptr,
                                       nullptr);
    else
        refill_wave<PRIV, false, false>(it.bvh.nodes, it.bvh.tris, it.bvh.node_count, it.rays,
                                        beg, end, refill, lds_stack + tid, (int)ls, (int)bs,
                                        wbits, it.hits.prim, it.hits.t, nullptr, nullptr,
                                        nullptr);
}

// ---------------------------------------------------------------------------
// BVH2 traversal of a SoftwareBVH (GpuBvhNode 32 B + GpuTriangle 48 B), the
// loop of benchmarks/exp0_probe/shaders/software_bvh_procedural.comp:119-171:
// the popped node's box is tested against the current closest t, a leaf
// tests its triangles in record order (|a| < 1e-8 rejected, accept
// t > tmin and t < closest), an inner node tests both children's boxes and
// pushes the hit ones far first.  Box test (:47-58): hit iff
// tFar >= max(0, tNear) and tNear < closest.  Same arithmetic order as the
// shader; the stack holds depth + 1 entries at most (LDS part + private).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool box2(const f4v& lo, const f4v& hi, float ox, float oy, float oz,
                                     float ix, float iy, float iz, float tcur, float& tnear) {
    const float t0x = (lo.x - ox) * ix, t1x = (hi.x - ox) * ix;
    const float t0y = (lo.y - oy) * iy, t1y = (hi.y - oy) * iy;
    const float t0z = (lo.z - oz) * iz, t1z = (hi.z - oz) * iz;
    const float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
    const float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
    tnear = tn;
    return tf >= fmaxf(0.0f, tn) && tn < tcur;
}

template <int PRIV, bool ANY, bool COUNT>
__global__ __launch_bounds__(kBlock) void k_trace_bvh2(
    const f4v* __restrict__ nodes, const f4v* __restrict__ tris, const uint32_t* __restrict__ order,
    uint32_t node_count, rt_rays_soa rays, uint32_t n, uint32_t lsu,
    uint32_t* __restrict__ prim_out, float* __restrict__ t_out, uint32_t* __restrict__ bits_out,
    uint32_t* __restrict__ box_out, uint32_t* __restrict__ tri_out) {
    extern __shared__ int32_t lds_stack[];  // dynamic: ls * blockDim.x entries
    const uint32_t tid = threadIdx.x;
    const int bs = (int)blockDim.x, ls = (int)lsu;
    int32_t* col = lds_stack + tid;
    const uint32_t r = blockIdx.x * blockDim.x + tid;
    const bool valid = r < n;
    float ox = 0.f, oy = 0.f, oz = 0.f, tmin = 1.f, dx = 0.f, dy = 0.f, dz = 1.f, tmax = 0.f;
    if (valid) {
        ox = rays.ox[r]; oy = rays.oy[r]; oz = rays.oz[r]; tmin = rays.tmin[r];
        dx = rays.dx[r]; dy = rays.dy[r]; dz = rays.dz[r]; tmax = rays.tmax[r];
    }
    // :108-110
    const float ix = 1.0f / (fabsf(dx) > 1e-8f ? dx : (dx >= 0.0f ? 1e-8f : -1e-8f));
    const float iy = 1.0f / (fabsf(dy) > 1e-8f ? dy : (dy >= 0.0f ? 1e-8f : -1e-8f));
    const float iz = 1.0f / (fabsf(dz) > 1e-8f ? dz : (dz >= 0.0f ? 1e-8f : -1e-8f));
    float closest = tmax;
    uint32_t prim = RT_MISS, boxT = 0, triT = 0;
    Stack<PRIV> st;
    int sp = 0, spo = 0;
    const int lso = (ls - 1) * bs;
    if (valid && node_count > 0) RT_SPUSH(0);
    bool done = false;
    while (sp > 0 && !done) {
        int32_t ni;
        RT_SPOP(ni);
        const f4v lo = nodes[2 * ni], hi = nodes[2 * ni + 1];
        if (COUNT) ++boxT;
        float tb;
        if (!box2(lo, hi, ox, oy, oz, ix, iy, iz, closest, tb)) continue;
        const int32_t left = __float_as_int(lo.w), right = __float_as_int(hi.w);
        if (left < 0) {
            const uint32_t first = (uint32_t)~left, cnt = (uint32_t)right;
            for (uint32_t i = 0; i < cnt; ++i) {
                if (COUNT) ++triT;
                const f4v* tp = tris + 3 * (size_t)(first + i);
                const f4v a = tp[0], b = tp[1], c = tp[2];
                float t;
                if (!mt_hit(a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, ox, oy, oz, dx, dy, dz,
                            &t, 1e-8f))
                    continue;
                if (t > tmin && t < closest) {
                    closest = t;
                    prim = order ? order[first + i] : first + i;
                    if (ANY) {
                        done = true;
                        break;
                    }
                }
            }
        } else {
            const f4v llo = nodes[2 * left], lhi = nodes[2 * left + 1];
            const f4v rlo = nodes[2 * right], rhi = nodes[2 * right + 1];
            if (COUNT) boxT += 2;
            float tl, tr;
            const bool hl = box2(llo, lhi, ox, oy, oz, ix, iy, iz, closest, tl);
            const bool hr = box2(rlo, rhi, ox, oy, oz, ix, iy, iz, closest, tr);
            if (hl && hr) {
                if (tl < tr) {
                    RT_SPUSH(right);
                    RT_SPUSH(left);
                } else {
                    RT_SPUSH(left);
                    RT_SPUSH(right);
                }
            } else if (hl) {
                RT_SPUSH(left);
            } else if (hr) {
                RT_SPUSH(right);
            }
        }
    }
    if (ANY) {
        write_wave_bits(bits_out, r - (tid & 63u), n, valid && prim != RT_MISS);
        if (valid && prim_out) prim_out[r] = prim;
    } else if (valid) {
        prim_out[r] = prim;
        t_out[r] = closest;
    }
    if (COUNT && valid) {
        box_out[r] = boxT;
        tri_out[r] = triT;
    }
}

struct TraceCfg {
    bool refill;
    bool packet;            // RT_SCHED_PACKET (occlusion bits only)
    uint32_t per_wave;      // multiple of 64, <= kMaxPerWave
    uint32_t refill_lanes;  // 1..64
    uint32_t block;         // 64, 128 or 256
    uint32_t lds;           // LDS stack entries per lane, kLdsMin..kLdsMax
    uint32_t order;         // any-hit child order (kOrderNear / Far / Slot)
    uint32_t grid_width;    // ray grid width for wave tiles; 0 = linear
    uint32_t tile_width;    // 8, 16 or 32 (with grid_width)
};

// Checks the launch-shape fields of opts (0 = default) into cfg; any_order
// and wave tiles only where the entry point offers them (allow: kWithOrder,
// kWithTiles).
constexpr uint32_t kWithOrder = 1u, kWithTiles = 2u;
bool shape_opts(const rt_trace_opts* o, TraceCfg& cfg, uint32_t allow = 0u) {
    if (!o) return true;
    if (o->block) cfg.block = o->block;
    if (o->lds_stack) cfg.lds = o->lds_stack;
    if (o->any_order && !(allow & kWithOrder)) return false;
    if ((o->tile_width || o->grid_width) && !(allow & kWithTiles)) return false;
    cfg.order = o->any_order;
    cfg.grid_width = o->grid_width;
    cfg.tile_width = o->tile_width;
    if (cfg.tile_width || cfg.grid_width) {
        if (cfg.tile_width != 8 && cfg.tile_width != 16 && cfg.tile_width != 32) return false;
        if (cfg.grid_width == 0 || cfg.grid_width % cfg.tile_width) return false;
    }
    return (cfg.block == 64 || cfg.block == 128 || cfg.block == 256) && cfg.lds >= kLdsMin &&
           cfg.lds <= kLdsMax && cfg.order <= (uint32_t)kOrderSlot;
}

size_t lds_bytes(const TraceCfg& cfg) { return (size_t)cfg.lds * cfg.block * 4; }

// Threads of a one-ray-per-thread launch over n rays: n, or with wave tiles
// 64 per tile of the grid (partial tiles at the bottom included); fills tm.
uint64_t launch_threads(const TraceCfg& cfg, uint32_t n, TileMap& tm) {
    tm = TileMap{cfg.grid_width, 0u, 0u};
    if (!cfg.tile_width) return n;
    tm.rows = n / cfg.grid_width;
    tm.tw_log2 = cfg.tile_width == 8 ? 3u : cfg.tile_width == 16 ? 4u : 5u;
    const uint64_t th = 64u >> tm.tw_log2;
    return (uint64_t)(cfg.grid_width / cfg.tile_width) * ((tm.rows + th - 1) / th) * 64u;
}

// Private overflow entries needed beyond the LDS part: at most 3 * depth + 1
// entries are ever live.
bool small_priv(uint32_t depth, uint32_t lds) {
    const uint32_t need = 3u * depth + 1u;
    return need <= lds + (uint32_t)kPrivSmall;
}

rt_status launch_packet(hipStream_t s, const rt_bvh_dev* b, const rt_rays_soa* rays,
                        uint32_t n, const TraceCfg& cfg, uint32_t* bits) {
    TileMap tm;
    const dim3 grid(ceil_div(launch_threads(cfg, n, tm), cfg.block)), block(cfg.block);
    const size_t shm = (size_t)(cfg.block / 64) * kPacketStack * sizeof(uint4);
    if (cfg.tile_width)
        hipLaunchKernelGGL(k_trace_packet<true>, grid, block, shm, s, b->nodes, b->tris,
                           b->node_count, *rays, n, bits, tm);
    else
        hipLaunchKernelGGL(k_trace_packet<false>, grid, block, shm, s, b->nodes, b->tris,
                           b->node_count, *rays, n, bits, tm);
    return launch_status("k_trace_packet");
}

template <int PRIV>
rt_status launch_v(hipStream_t s, const rt_bvh_dev* b, const rt_rays_soa* rays, uint32_t n,
                   uint32_t flags, const TraceCfg& cfg, rt_hits_soa* hits, uint32_t* bits,
                   uint32_t* box, uint32_t* tri) {
    const bool any = flags & RT_ANY;
    const bool cnt = flags & RT_COUNTERS;
    uint32_t* prim = hits ? hits->prim : nullptr;
    float* t = hits ? hits->t : nullptr;
    const size_t shm = lds_bytes(cfg);
    if (cfg.refill && any && !cnt && !prim) {
        // occlusion bits only: the bits-only node step; waves x per_wave
        // covers the launch's slots (n, or 64 per wave tile)
        TileMap tm;
        const uint64_t slots = launch_threads(cfg, n, tm);
        if (slots > 0xFFFFFFFFull) return RT_ERR_CAPACITY;
        const uint32_t blocks = ceil_div(slots, (uint64_t)cfg.per_wave * (cfg.block / 64));
        const size_t shm_bits = shm + (size_t)(cfg.block / 64) * (cfg.per_wave / 32) * 4;
        if (cfg.tile_width)
            hipLaunchKernelGGL((k_trace_refill_bits<PRIV, true>), dim3(blocks), dim3(cfg.block),
                               shm_bits, s, b->nodes, b->tris, b->node_count, *rays, n,
                               (uint32_t)slots, cfg.per_wave, cfg.refill_lanes, cfg.lds, bits, tm);
        else
            hipLaunchKernelGGL((k_trace_refill_bits<PRIV, false>), dim3(blocks), dim3(cfg.block),
                               shm_bits, s, b->nodes, b->tris, b->node_count, *rays, n,
                               (uint32_t)slots, cfg.per_wave, cfg.refill_lanes, cfg.lds, bits, tm);
        return launch_status("k_trace_refill_bits");
    }
    if (cfg.refill) {
        // waves x per_wave covers [0, n); per_wave is a multiple of 64 and at
        // most kMaxPerWave, so each wave owns whole occlusion words
        const uint32_t blocks = ceil_div(n, (uint64_t)cfg.per_wave * (cfg.block / 64));
#define RT_P(A, C)                                                                            \
    hipLaunchKernelGGL((k_trace_persist<PRIV, A, C>), dim3(blocks), dim3(cfg.block), shm, s,  \
                       b->nodes, b->tris, b->node_count, *rays, n, cfg.per_wave,              \
                       cfg.refill_lanes, cfg.lds, prim, t, bits, box, tri)
        if (any && cnt) RT_P(true, true);
        else if (any) RT_P(true, false);
        else if (cnt) RT_P(false, true);
        else RT_P(false, false);
#undef RT_P
        return launch_status("k_trace_persist");
    }
    TileMap tm;
    dim3 grid(ceil_div(launch_threads(cfg, n, tm), cfg.block)), block(cfg.block);
    // occlusion bits only (no counters, no hit ids): the order does not change
    // the bits, so near-first is run as kOrderInner
    const bool bits_only = any && !cnt && !prim;
    if (b->format == RT_NODES_Q8) {
#define RT_Q(A, C, O, T)                                                                      \
    hipLaunchKernelGGL((k_trace<PRIV, A, C, O, T, true>), grid, block, shm, s, b->nodes,      \
                       b->tris, b->node_count, *rays, n, cfg.lds, prim, t, bits, box, tri, tm)
        if (cfg.tile_width) {
            if (cfg.order == kOrderFar) RT_Q(true, false, kOrderFar, true);
            else if (cfg.order == kOrderSlot) RT_Q(true, false, kOrderSlot, true);
            else if (bits_only) RT_Q(true, false, kOrderInner, true);
            else if (any && cnt) RT_Q(true, true, kOrderNear, true);
            else if (any) RT_Q(true, false, kOrderNear, true);
            else if (cnt) RT_Q(false, true, kOrderNear, true);
            else RT_Q(false, false, kOrderNear, true);
        } else {
            if (cfg.order == kOrderFar) RT_Q(true, false, kOrderFar, false);
            else if (cfg.order == kOrderSlot) RT_Q(true, false, kOrderSlot, false);
            else if (bits_only) RT_Q(true, false, kOrderInner, false);
            else if (any && cnt) RT_Q(true, true, kOrderNear, false);
            else if (any) RT_Q(true, false, kOrderNear, false);
            else if (cnt) RT_Q(false, true, kOrderNear, false);
            else RT_Q(false, false, kOrderNear, false);
        }
#undef RT_Q
        return launch_status("k_trace");
    }
#define RT_L(A, C, O, T)                                                                      \
    hipLaunchKernelGGL((k_trace<PRIV, A, C, O, T>), grid, block, shm, s, b->nodes, b->tris,   \
                       b->node_count, *rays, n, cfg.lds, prim, t, bits, box, tri, tm)
    if (cfg.tile_width) {
        if (cfg.order == kOrderFar) RT_L(true, false, kOrderFar, true);
        else if (cfg.order == kOrderSlot) RT_L(true, false, kOrderSlot, true);
        else if (bits_only) RT_L(true, false, kOrderInner, true);
        else if (any && cnt) RT_L(true, true, kOrderNear, true);
        else if (any) RT_L(true, false, kOrderNear, true);
        else if (cnt) RT_L(false, true, kOrderNear, true);
        else RT_L(false, false, kOrderNear, true);
    } else {
        if (cfg.order == kOrderFar) RT_L(true, false, kOrderFar, false);
        else if (cfg.order == kOrderSlot) RT_L(true, false, kOrderSlot, false);
        else if (bits_only) RT_L(true, false, kOrderInner, false);
        else if (any && cnt) RT_L(true, true, kOrderNear, false);
        else if (any) RT_L(true, false, kOrderNear, false);
        else if (cnt) RT_L(false, true, kOrderNear, false);
        else RT_L(false, false, kOrderNear, false);
    }
#undef RT_L
    return launch_status("k_trace");
}

}  // namespace
}  // namespace rt

using namespace rt;

extern "C" rt_status rt_trace_ex(rt_stream stream, const rt_bvh_dev* bvh, const rt_rays_soa* rays,
                                 uint32_t n, uint32_t flags, const rt_trace_opts* opts,
                                 rt_hits_soa* hits, uint32_t* occl_bits, uint32_t* box_tests,
                                 uint32_t* tri_tests) {
    if (!bvh || !rays) return RT_ERR_ARG;
    if (flags & ~(RT_ANY | RT_COUNTERS | RT_TRACE_PERSISTENT)) return RT_ERR_ARG;
    TraceCfg cfg{(flags & RT_TRACE_PERSISTENT) != 0, false, 1024u, 16u, (uint32_t)kBlock,
                 kLdsStack};
    if (opts) {
        if (opts->schedule > RT_SCHED_PACKET) return RT_ERR_ARG;
        cfg.refill = cfg.refill || opts->schedule == RT_SCHED_REFILL;
        cfg.packet = opts->schedule == RT_SCHED_PACKET;
        if (opts->rays_per_wave) cfg.per_wave = opts->rays_per_wave;
        if (opts->refill_lanes) cfg.refill_lanes = opts->refill_lanes;
    }
    if (!shape_opts(opts, cfg, kWithOrder | kWithTiles)) return RT_ERR_ARG;
    if (cfg.per_wave % 64 || cfg.per_wave > kMaxPerWave || cfg.refill_lanes > 64)
        return RT_ERR_ARG;
    // another child order: occlusion bits only (no counters, no hit ids), one
    // ray per thread
    if (cfg.order != kOrderNear &&
        (!(flags & RT_ANY) || (flags & RT_COUNTERS) || (hits && hits->prim) || cfg.refill))
        return RT_ERR_ARG;
    // packets: occlusion bits only, their own order
    if (cfg.packet && (!(flags & RT_ANY) || (flags & RT_COUNTERS) || (hits && hits->prim) ||
                       cfg.refill || cfg.order != kOrderNear))
        return RT_ERR_ARG;
    // wave tiles: whole rows; one ray per thread, or refill for occlusion bits only
    const bool bits_only = (flags & RT_ANY) && !(flags & RT_COUNTERS) && !(hits && hits->prim);
    if (cfg.tile_width && ((cfg.refill && !bits_only) || n % cfg.grid_width)) return RT_ERR_ARG;
    if (n == 0) return RT_OK;
    if (!rays->ox || !rays->oy || !rays->oz || !rays->tmin || !rays->dx || !rays->dy ||
        !rays->dz || !rays->tmax)
        return RT_ERR_ARG;
    const bool any = flags & RT_ANY;
    if (any && !occl_bits) return RT_ERR_ARG;
    if (!any && (!hits || !hits->prim || !hits->t)) return RT_ERR_ARG;
    if ((flags & RT_COUNTERS) && (!box_tests || !tri_tests)) return RT_ERR_ARG;
    if (bvh->node_count > 0 && (!bvh->nodes || !bvh->tris)) return RT_ERR_ARG;
    if (bvh->format > RT_NODES_Q8) return RT_ERR_ARG;
    if (bvh->format == RT_NODES_Q8 && (cfg.refill || cfg.packet)) return RT_ERR_UNSUPPORTED;
    // depth 0 = unknown (<= kMaxDepth)
    const uint32_t depth = bvh->depth ? bvh->depth : kMaxDepth;
    if (depth > kMaxDepth) return RT_ERR_CAPACITY;
    hipStream_t s = as_stream(stream);
    if (cfg.packet) return launch_packet(s, bvh, rays, n, cfg, occl_bits);
    if (small_priv(depth, cfg.lds))
        return launch_v<kPrivSmall>(s, bvh, rays, n, flags, cfg, hits, occl_bits, box_tests,
                                    tri_tests);
    return launch_v<kPrivLarge>(s, bvh, rays, n, flags, cfg, hits, occl_bits, box_tests,
                                tri_tests);
}

extern "C" rt_status rt_prim_records(rt_stream stream, const rt_bvh_dev* bvh, uint32_t ids,
                                     uint32_t* d_record) {
    if (!bvh) return RT_ERR_ARG;
    if (ids == 0) return RT_OK;
    if (!d_record || (bvh->tri_count > 0 && !bvh->tris)) return RT_ERR_ARG;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_fill_u32, dim3(ceil_div(ids, 256)), dim3(256), 0, s, d_record, ids,
                       RT_MISS);
    if (bvh->tri_count > 0)
        hipLaunchKernelGGL(k_prim_records, dim3(ceil_div(bvh->tri_count, 256)), dim3(256), 0, s,
                           bvh->tris, bvh->tri_count, ids, d_record);
    return launch_status("rt_prim_records");
}

namespace rt {
namespace {
// The hinted trace's launch shape and scratch layout; false on bad arguments.
struct HintLayout {
    TraceCfg cfg;
    TileMap tm;
    uint64_t slots;
    uint32_t nseg, seg_words;
    size_t open_bytes, total;
};
bool hint_layout(uint32_t n, const rt_trace_opts* opts, const rt_occluder_hint* hint,
                 HintLayout& L) {
    L.cfg = TraceCfg{true, false, 1024u, 16u, (uint32_t)kBlock, kLdsStack};
    if (opts) {
        if (opts->schedule != RT_SCHED_SIMPLE && opts->schedule != RT_SCHED_REFILL) return false;
        if (opts->rays_per_wave) L.cfg.per_wave = opts->rays_per_wave;
        if (opts->refill_lanes) L.cfg.refill_lanes = opts->refill_lanes;
    }
    if (!shape_opts(opts, L.cfg, kWithTiles)) return false;
    if (L.cfg.per_wave == 0 || L.cfg.per_wave % 64 || L.cfg.per_wave > kMaxPerWave ||
        L.cfg.refill_lanes == 0 || L.cfg.refill_lanes > 64)
        return false;
    if (L.cfg.tile_width && n % L.cfg.grid_width) return false;
    if (hint->flags & ~RT_HINT_WAVE) return false;
    L.slots = launch_threads(L.cfg, n, L.tm);
    if (L.slots > 0xFFFFFFFFull - kMaxPerWave) return false;
    L.seg_words = L.cfg.per_wave / 32;
    L.nseg = ceil_div(L.slots, L.cfg.per_wave);
    // the open mask covers whole segments (the last one padded with closed
    // slots), then the segment counter
    L.open_bytes = ((size_t)L.nseg * L.seg_words * 4 + 255) & ~(size_t)255;
    L.total = L.open_bytes + 256;
    return true;
}
}  // namespace
}  // namespace rt

extern "C" rt_status rt_trace_occluded_hinted_scratch_bytes(uint32_t n,
                                                            const rt_trace_opts* opts,
                                                            const rt_occluder_hint* hint,
                                                            size_t* bytes) {
    if (!hint || !bytes) return RT_ERR_ARG;
    HintLayout L;
    if (!hint_layout(n, opts, hint, L)) return RT_ERR_ARG;
    *bytes = L.total;
    return RT_OK;
}

extern "C" rt_status rt_trace_occluded_hinted(rt_stream stream, const rt_bvh_dev* bvh,
                                              const rt_rays_soa* rays, uint32_t n,
                                              const rt_trace_opts* opts,
                                              const rt_occluder_hint* hint, uint32_t* occl_bits,
                                              void* d_scratch, size_t scratch_bytes) {
    if (!bvh || !rays || !hint) return RT_ERR_ARG;
    HintLayout L;
    if (!hint_layout(n, opts, hint, L)) return RT_ERR_ARG;
    if (n == 0) return RT_OK;
    if (!rays->ox || !rays->oy || !rays->oz || !rays->tmin || !rays->dx || !rays->dy ||
        !rays->dz || !rays->tmax || !occl_bits || !hint->prim || (hint->ids && !hint->record))
        return RT_ERR_ARG;
    if (bvh->node_count > 0 && (!bvh->nodes || !bvh->tris)) return RT_ERR_ARG;
    if (bvh->format != RT_NODES_F32) return RT_ERR_UNSUPPORTED;
    const uint32_t depth = bvh->depth ? bvh->depth : kMaxDepth;
    if (depth > kMaxDepth) return RT_ERR_CAPACITY;
    if (!d_scratch || scratch_bytes < L.total) return RT_ERR_CAPACITY;
    const TraceCfg& cfg = L.cfg;
    char* sc = static_cast<char*>(d_scratch);
    // layers of `period` rays: the die bands of queue_segment when whole
    // segments (and, with wave tiles, whole tile rows) make up each layer
    const uint32_t period = hint->period ? hint->period : n;
    const uint32_t tile_rows = cfg.tile_width ? 64u / cfg.tile_width : 1u;
    const bool layered = n % period == 0 && period % cfg.per_wave == 0 &&
                         (!cfg.tile_width || period % (cfg.grid_width * tile_rows) == 0) &&
                         (uint64_t)n == L.slots;
    const HintPlan pl{reinterpret_cast<uint32_t*>(sc), reinterpret_cast<uint32_t*>(sc + L.open_bytes),
                      L.nseg, L.seg_words, layered ? period / cfg.per_wave : 0u,
                      layered ? n / period : 0u};
    const HintArgs h{hint->prim, hint->record, period, hint->ids,
                     (hint->flags & RT_HINT_WAVE) ? 1u : 0u};
    hipStream_t s = as_stream(stream);
    const TileMap tm = L.tm;
    // pass 1 over every slot of whole segments (k_hint_pass closes the slots
    // past the launch's)
    const uint64_t pass_slots = (uint64_t)L.nseg * cfg.per_wave;
    const dim3 g1(ceil_div(pass_slots, kBlock));
    if (cfg.tile_width)
        hipLaunchKernelGGL(k_hint_pass<true>, g1, dim3(kBlock), 0, s, bvh->tris, bvh->node_count,
                           bvh->tri_count, *rays, n, (uint32_t)L.slots, occl_bits, tm, h, pl);
    else
        hipLaunchKernelGGL(k_hint_pass<false>, g1, dim3(kBlock), 0, s, bvh->tris,
                           bvh->node_count, bvh->tri_count, *rays, n, (uint32_t)L.slots,
                           occl_bits, tm, h, pl);
    RT_TRY_HIP(hipGetLastError());
    // the trace pass, persistent: as many workgroups as fit on the device at once
    const size_t shm = lds_bytes(cfg) + (size_t)(cfg.block / 64) * kRing * 4;
    const void* fn;
    if (small_priv(depth, cfg.lds))
        fn = cfg.tile_width ? reinterpret_cast<const void*>(&k_hint_trace<kPrivSmall, true>)
                            : reinterpret_cast<const void*>(&k_hint_trace<kPrivSmall, false>);
    else
        fn = cfg.tile_width ? reinterpret_cast<const void*>(&k_hint_trace<kPrivLarge, true>)
                            : reinterpret_cast<const void*>(&k_hint_trace<kPrivLarge, false>);
    int dev = 0, cus = 0, per_cu = 0;
    RT_TRY_HIP(hipGetDevice(&dev));
    RT_TRY_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    RT_TRY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, (int)cfg.block, shm));
    const uint64_t fit = (uint64_t)(per_cu > 0 ? per_cu : 1) * (uint64_t)(cus > 0 ? cus : 1);
    const uint32_t blocks = (uint32_t)(fit < L.nseg ? fit : L.nseg);
    const rt_node4* nodes = bvh->nodes;
    const rt_tri64* tris = bvh->tris;
    rt_rays_soa rs = *rays;
    TileMap tmv = tm;
    HintPlan plv = pl;
    uint32_t refill = cfg.refill_lanes, lsu = cfg.lds;
    uint32_t* bits = occl_bits;
    void* args[] = {&nodes, &tris, &rs, &tmv, &plv, &refill, &lsu, &bits};
    RT_TRY_HIP(hipLaunchKernel(fn, dim3(blocks), dim3(cfg.block), args, shm, s));
    return launch_status("rt_trace_occluded_hinted");
}

extern "C" rt_status rt_radv_describe(const void* host_image, uint64_t bytes,
                                     rt_radv_info* info) {
    if (!host_image || !info) return RT_ERR_ARG;
    if (bytes < 136) return RT_ERR_ARG;
    const uint8_t* b = static_cast<const uint8_t*>(host_image);
    rt_radv_info o{};
    const auto g32 = [&](size_t off) { uint32_t v; std::memcpy(&v, b + off, 4); return v; };
    const auto g64 = [&](size_t off) { uint64_t v; std::memcpy(&v, b + off, 8); return v; };
    // include/radv_bvh.hpp:59-82 RadvAccelStructHeader field offsets
    o.bvh_offset = g32(0x00);
    o.root_flags = g32(0x04);
    std::memcpy(o.aabb, b + 0x08, 24);
    for (int i = 0; i < 3; ++i) o.update_dispatch_size[i] = g32(0x20 + 4 * i);
    o.compacted_size = g64(0x30);
    o.serialization_size = g64(0x38);
    for (int i = 0; i < 3; ++i) o.copy_dispatch_size[i] = g32(0x40 + 4 * i);
    o.size = g64(0x50);
    o.geometry_type = g32(0x58);
    o.geometry_count = g32(0x5C);
    o.leaf_nodes_offset = g64(0x60);
    o.primitive_count = g64(0x68);
    o.instance_count = g64(0x70);
    o.leaf_node_offsets_offset = g32(0x78);
    o.build_flags = g32(0x7C);
    o.type = g32(0x80);
    *info = o;
    if (o.bvh_offset % 64 || o.leaf_nodes_offset < o.bvh_offset ||
        (o.leaf_nodes_offset - o.bvh_offset) % 128 ||
        (o.leaf_nodes_offset - o.bvh_offset) / 128 > 0xFFFFFFFFull ||
        o.primitive_count > 0xFFFFFFFFull ||
        o.leaf_nodes_offset + 64 * o.primitive_count > bytes)
        return RT_ERR_ARG;
    info->box_count = (uint32_t)((o.leaf_nodes_offset - o.bvh_offset) / 128);
    if (o.bvh_offset + 128 <= bytes && info->box_count > 0)
        for (int i = 0; i < 4; ++i) info->root_children[i] = g32(o.bvh_offset + 4 * i);
    return RT_OK;
}

extern "C" rt_status rt_trace_radv(rt_stream stream, const rt_radv_dev* img,
                                   const rt_rays_soa* rays, uint32_t n, uint32_t flags,
                                   const rt_trace_opts* opts, rt_hits_soa* hits,
                                   uint32_t* occl_bits, uint32_t* box_tests,
                                   uint32_t* tri_tests) {
    if (!img || !rays) return RT_ERR_ARG;
    if (flags & ~(RT_ANY | RT_COUNTERS)) return RT_ERR_ARG;
    TraceCfg cfg{false, false, 0u, 0u, (uint32_t)kBlock, kLdsStack};
    if (opts && opts->schedule != RT_SCHED_SIMPLE) return RT_ERR_ARG;
    if (!shape_opts(opts, cfg)) return RT_ERR_ARG;
    // the sections must lie inside the image, 16-byte aligned for the loads
    const uint64_t leaf = (uint64_t)img->bvh_offset + 128ull * img->box_count;
    if (leaf + 64ull * img->tri_count > img->bytes) return RT_ERR_CAPACITY;
    if (img->box_count && (!img->image || img->bvh_offset % 64 ||
                           reinterpret_cast<uintptr_t>(img->image) % 16))
        return RT_ERR_ARG;
    if (n == 0) return RT_OK;
    if (!rays->ox || !rays->oy || !rays->oz || !rays->tmin || !rays->dx || !rays->dy ||
        !rays->dz || !rays->tmax)
        return RT_ERR_ARG;
    const bool any = flags & RT_ANY, cnt = flags & RT_COUNTERS;
    if (any && !occl_bits) return RT_ERR_ARG;
    if (!any && (!hits || !hits->prim || !hits->t)) return RT_ERR_ARG;
    if (cnt && (!box_tests || !tri_tests)) return RT_ERR_ARG;
    const uint32_t depth = img->depth ? img->depth : kMaxDepth;
    if (depth > kMaxDepth) return RT_ERR_CAPACITY;
    const uint8_t* base = static_cast<const uint8_t*>(img->image);
    const rt_node4* nd = reinterpret_cast<const rt_node4*>(base + img->bvh_offset);
    const rt_tri64* tr = reinterpret_cast<const rt_tri64*>(base + leaf);
    uint32_t* prim = hits ? hits->prim : nullptr;
    float* t = hits ? hits->t : nullptr;
    const dim3 grid(ceil_div(n, cfg.block)), block(cfg.block);
    const size_t shm = lds_bytes(cfg);
    hipStream_t s = as_stream(stream);
#define RT_R(P, A, C)                                                                          \
    hipLaunchKernelGGL((k_trace_radv<P, A, C>), grid, block, shm, s, nd, tr, img->box_count,   \
                       img->tri_count, *rays, n, cfg.lds, prim, t, occl_bits, box_tests,       \
                       tri_tests)
    if (small_priv(depth, cfg.lds)) {
        if (any && cnt) RT_R(kPrivSmall, true, true);
        else if (any) RT_R(kPrivSmall, true, false);
        else if (cnt) RT_R(kPrivSmall, false, true);
        else RT_R(kPrivSmall, false, false);
    } else {
        if (any && cnt) RT_R(kPrivLarge, true, true);
        else if (any) RT_R(kPrivLarge, true, false);
        else if (cnt) RT_R(kPrivLarge, false, true);
        else RT_R(kPrivLarge, false, false);
    }
#undef RT_R
    return launch_status("k_trace_radv");
}

extern "C" rt_status rt_bvh2_depth(const void* nodes, uint32_t node_count, uint32_t* depth) {
    if (!depth || (node_count && !nodes)) return RT_ERR_ARG;
    *depth = 0;
    if (node_count == 0) return RT_OK;
    struct N2 {
        float mn[3];
        int32_t l;
        float mx[3];
        int32_t r;
    };
    const N2* nd = static_cast<const N2*>(nodes);
    std::vector<std::pair<uint32_t, uint32_t>> st{{0u, 1u}};
    uint32_t deep = 0, visited = 0;
    while (!st.empty()) {
        const auto [i, d] = st.back();
        st.pop_back();
        if (i >= node_count || ++visited > node_count) return RT_ERR_ARG;  // not a tree
        deep = d > deep ? d : deep;
        if (nd[i].l >= 0) {
            st.push_back({(uint32_t)nd[i].l, d + 1});
            st.push_back({(uint32_t)nd[i].r, d + 1});
        }
    }
    *depth = deep;
    return RT_OK;
}

extern "C" rt_status rt_trace_bvh2(rt_stream stream, const rt_bvh2_dev* bvh,
                                   const rt_rays_soa* rays, uint32_t n, uint32_t flags,
                                   const rt_trace_opts* opts, rt_hits_soa* hits,
                                   uint32_t* occl_bits, uint32_t* box_tests,
                                   uint32_t* tri_tests) {
    if (!bvh || !rays) return RT_ERR_ARG;
    if (flags & ~(RT_ANY | RT_COUNTERS)) return RT_ERR_ARG;
    TraceCfg cfg{false, false, 0u, 0u, (uint32_t)kBlock, kLdsStack};
    if (opts && opts->schedule != RT_SCHED_SIMPLE) return RT_ERR_ARG;
    if (!shape_opts(opts, cfg)) return RT_ERR_ARG;
    if (n == 0) return RT_OK;
    if (!rays->ox || !rays->oy || !rays->oz || !rays->tmin || !rays->dx || !rays->dy ||
        !rays->dz || !rays->tmax)
        return RT_ERR_ARG;
    const bool any = flags & RT_ANY, cnt = flags & RT_COUNTERS;
    if (any && !occl_bits) return RT_ERR_ARG;
    if (!any && (!hits || !hits->prim || !hits->t)) return RT_ERR_ARG;
    if (cnt && (!box_tests || !tri_tests)) return RT_ERR_ARG;
    if (bvh->node_count > 0 && (!bvh->nodes || !bvh->tris48)) return RT_ERR_ARG;
    // a BVH2 stack holds at most depth + 1 entries; depth 0 = unknown (<= 64)
    const uint32_t depth = bvh->depth ? bvh->depth : kMaxDepth;
    if (depth > kMaxDepth) return RT_ERR_CAPACITY;
    const bool small = depth + 1u <= cfg.lds + (uint32_t)kPrivSmall;
    uint32_t* prim = hits ? hits->prim : nullptr;
    float* t = hits ? hits->t : nullptr;
    const f4v* nd = static_cast<const f4v*>(bvh->nodes);
    const f4v* tr = reinterpret_cast<const f4v*>(bvh->tris48);
    const dim3 grid(ceil_div(n, cfg.block)), block(cfg.block);
    const size_t shm = lds_bytes(cfg);
    hipStream_t s = as_stream(stream);
#define RT_B2(P, A, C)                                                                         \
    hipLaunchKernelGGL((k_trace_bvh2<P, A, C>), grid, block, shm, s, nd, tr, bvh->order,       \
                       bvh->node_count, *rays, n, cfg.lds, prim, t, occl_bits, box_tests,      \
                       tri_tests)
    if (small) {
        if (any && cnt) RT_B2(kPrivSmall, true, true);
        else if (any) RT_B2(kPrivSmall, true, false);
        else if (cnt) RT_B2(kPrivSmall, false, true);
        else RT_B2(kPrivSmall, false, false);
    } else {
        if (any && cnt) RT_B2(kPrivLarge, true, true);
        else if (any) RT_B2(kPrivLarge, true, false);
        else if (cnt) RT_B2(kPrivLarge, false, true);
        else RT_B2(kPrivLarge, false, false);
    }
#undef RT_B2
    return launch_status("k_trace_bvh2");
}

extern "C" rt_status rt_trace(rt_stream stream, const rt_bvh_dev* bvh, const rt_rays_soa* rays,
                              uint32_t n, uint32_t flags, rt_hits_soa* hits, uint32_t* occl_bits,
                              uint32_t* box_tests, uint32_t* tri_tests) {
    return rt_trace_ex(stream, bvh, rays, n, flags, nullptr, hits, occl_bits, box_tests,
                       tri_tests);
}

extern "C" rt_status rt_trace_batch_plan_ex(const rt_trace_item* items, uint32_t count,
                                            uint32_t chunk_rays, uint32_t* chunk_start,
                                            uint32_t* depth) {
    if ((count && !items) || !chunk_start) return RT_ERR_ARG;
    if (chunk_rays == 0 || chunk_rays % 64 || chunk_rays > kMaxPerWave) return RT_ERR_ARG;
    uint64_t c = 0;
    uint32_t deep = 1;
    for (uint32_t i = 0; i < count; ++i) {
        const rt_trace_item& it = items[i];
        if (it.flags & ~RT_ANY) return RT_ERR_ARG;
        if (it.count) {
            const rt_rays_soa& r = it.rays;
            if (!r.ox || !r.oy || !r.oz || !r.tmin || !r.dx || !r.dy || !r.dz || !r.tmax)
                return RT_ERR_ARG;
            if ((it.flags & RT_ANY) ? !it.occlusion_bits : (!it.hits.prim || !it.hits.t))
                return RT_ERR_ARG;
            if (it.bvh.node_count > 0 && (!it.bvh.nodes || !it.bvh.tris)) return RT_ERR_ARG;
        }
        if (it.bvh.format != RT_NODES_F32) return RT_ERR_UNSUPPORTED;
        const uint32_t d = it.bvh.depth ? it.bvh.depth : kMaxDepth;
        if (d > kMaxDepth) return RT_ERR_CAPACITY;
        deep = d > deep ? d : deep;
        chunk_start[i] = (uint32_t)c;
        c += ceil_div(it.count, chunk_rays);
        if (c > 0xFFFFFFFFull / kBlock) return RT_ERR_CAPACITY;
    }
    chunk_start[count] = (uint32_t)c;
    if (depth) *depth = deep;
    return RT_OK;
}

extern "C" rt_status rt_trace_batch_plan(const rt_trace_item* items, uint32_t count,
                                         uint32_t* chunk_start, uint32_t* depth) {
    return rt_trace_batch_plan_ex(items, count, 64u, chunk_start, depth);
}

extern "C" rt_status rt_trace_batch_ex(rt_stream stream, const rt_trace_item* d_items,
                                       const uint32_t* d_chunk_start, uint32_t count,
                                       uint32_t total_chunks, uint32_t depth,
                                       const rt_trace_opts* opts) {
    TraceCfg cfg{false, false, 1024u, 16u, (uint32_t)kBlock, kLdsStack};
    if (opts) {
        if (opts->schedule > RT_SCHED_REFILL) return RT_ERR_ARG;
        cfg.refill = opts->schedule == RT_SCHED_REFILL;
        if (opts->rays_per_wave) cfg.per_wave = opts->rays_per_wave;
        if (opts->refill_lanes) cfg.refill_lanes = opts->refill_lanes;
    }
    if (!shape_opts(opts, cfg)) return RT_ERR_ARG;
    if (cfg.per_wave % 64 || cfg.per_wave > kMaxPerWave || cfg.refill_lanes > 64)
        return RT_ERR_ARG;
    if (total_chunks == 0) return RT_OK;
    if (!d_items || !d_chunk_start || count == 0) return RT_ERR_ARG;
    if (depth == 0 || depth > kMaxDepth) return RT_ERR_ARG;
    hipStream_t s = as_stream(stream);
    const dim3 grid(ceil_div(total_chunks, cfg.block / 64)), block(cfg.block);
    const size_t shm = lds_bytes(cfg);
    if (cfg.refill) {
        if (small_priv(depth, cfg.lds))
            hipLaunchKernelGGL(k_trace_batch_refill<kPrivSmall>, grid, block, shm, s, d_items,
                               d_chunk_start, count, total_chunks, cfg.per_wave,
                               cfg.refill_lanes, cfg.lds);
        else
            hipLaunchKernelGGL(k_trace_batch_refill<kPrivLarge>, grid, block, shm, s, d_items,
                               d_chunk_start, count, total_chunks, cfg.per_wave,
                               cfg.refill_lanes, cfg.lds);
        return launch_status("k_trace_batch_refill");
    }
    if (small_priv(depth, cfg.lds))
        hipLaunchKernelGGL(k_trace_batch<kPrivSmall>, grid, block, shm, s, d_items, d_chunk_start,
                           count, total_chunks, cfg.lds);
    else
        hipLaunchKernelGGL(k_trace_batch<kPrivLarge>, grid, block, shm, s, d_items, d_chunk_start,
                           count, total_chunks, cfg.lds);
    return launch_status("k_trace_batch");
}

namespace {
// Byte layout of a submit table (staging and device copy alike): items, the
// optional payload generators, the chunk table; each part 256-B aligned.
struct SubmitLayout {
    size_t items = 0, gen = 0, chunk = 0, total = 0;
};
SubmitLayout submit_layout(uint32_t count, bool with_gen) {
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    SubmitLayout L;
    size_t off = al((size_t)count * sizeof(rt_trace_item));
    if (with_gen) {
        L.gen = off;
        off += al((size_t)count * sizeof(rt_ray_gen));
    }
    L.chunk = off;
    off += al(((size_t)count + 1) * 4);
    L.total = off;
    return L;
}
}  // namespace

extern "C" rt_status rt_trace_batch_submit_bytes(uint32_t count, int with_gen, size_t* bytes) {
    if (!bytes) return RT_ERR_ARG;
    *bytes = submit_layout(count, with_gen != 0).total;
    return RT_OK;
}

extern "C" rt_status rt_trace_batch_submit(rt_stream stream, const rt_trace_item* items,
                                           const rt_ray_gen* gen, uint32_t count,
                                           const rt_trace_opts* opts, void* pinned,
                                           size_t pinned_bytes, void* d_table,
                                           size_t table_bytes) {
    if (count == 0) return RT_OK;
    if (!items || !pinned || !d_table) return RT_ERR_ARG;
    if (reinterpret_cast<uintptr_t>(d_table) % 256) return RT_ERR_ARG;
    const SubmitLayout L = submit_layout(count, gen != nullptr);
    if (pinned_bytes < L.total || table_bytes < L.total) return RT_ERR_CAPACITY;
    uint32_t chunk_rays = 64u;
    if (opts && opts->schedule == RT_SCHED_REFILL)
        chunk_rays = opts->rays_per_wave ? opts->rays_per_wave : 1024u;
    char* h = static_cast<char*>(pinned);
    uint32_t* h_chunk = reinterpret_cast<uint32_t*>(h + L.chunk);
    uint32_t depth = 0;
    rt_status st = rt_trace_batch_plan_ex(items, count, chunk_rays, h_chunk, &depth);
    if (st != RT_OK) return st;
    std::memcpy(h, items, (size_t)count * sizeof(rt_trace_item));
    if (gen) std::memcpy(h + L.gen, gen, (size_t)count * sizeof(rt_ray_gen));
    const uint32_t total = h_chunk[count];
    hipStream_t s = as_stream(stream);
    RT_TRY_HIP(hipMemcpyAsync(d_table, pinned, L.total, hipMemcpyHostToDevice, s));
    char* d = static_cast<char*>(d_table);
    const auto d_items = reinterpret_cast<const rt_trace_item*>(d);
    const auto d_chunk = reinterpret_cast<const uint32_t*>(d + L.chunk);
    if (gen) {
        st = launch_ray_gen(s, d_items, reinterpret_cast<const rt_ray_gen*>(d + L.gen), d_chunk,
                            count, total, chunk_rays);
        if (st != RT_OK) return st;
    }
    return rt_trace_batch_ex(stream, d_items, d_chunk, count, total, depth, opts);
}

extern "C" rt_status rt_trace_batch(rt_stream stream, const rt_trace_item* d_items,
                                    const uint32_t* d_chunk_start, uint32_t count,
                                    uint32_t total_chunks, uint32_t depth) {
    return rt_trace_batch_ex(stream, d_items, d_chunk_start, count, total_chunks, depth, nullptr);
}

extern "C" rt_status rt_trace_scene_ex(rt_stream stream, const rt_instance* d_instances,
                                       uint32_t count, uint32_t max_depth,
                                       const rt_rays_soa* rays, uint32_t n, uint32_t flags,
                                       const rt_trace_opts* opts, rt_hits_soa* hits,
                                       uint32_t* d_inst, uint32_t* occl_bits) {
    TraceCfg cfg{false, false, 0u, 0u, (uint32_t)kBlock, kLdsStack};
    if (opts && opts->schedule != RT_SCHED_SIMPLE) return RT_ERR_ARG;
    if (!shape_opts(opts, cfg, kWithTiles)) return RT_ERR_ARG;
    if (!rays || (flags & ~RT_ANY)) return RT_ERR_ARG;
    if (cfg.tile_width && n % cfg.grid_width) return RT_ERR_ARG;  // whole rows
    if (n == 0) return RT_OK;
    if ((count && !d_instances) || !rays->ox || !rays->oy || !rays->oz || !rays->tmin ||
        !rays->dx || !rays->dy || !rays->dz || !rays->tmax)
        return RT_ERR_ARG;
    const bool any = flags & RT_ANY;
    if (any && !occl_bits) return RT_ERR_ARG;
    if (!any && (!hits || !hits->prim || !hits->t)) return RT_ERR_ARG;
    const uint32_t depth = max_depth ? max_depth : kMaxDepth;
    if (depth > kMaxDepth) return RT_ERR_CAPACITY;
    uint32_t* prim = hits ? hits->prim : nullptr;
    float* t = hits ? hits->t : nullptr;
    TileMap tm;
    const dim3 grid(ceil_div(launch_threads(cfg, n, tm), cfg.block)), block(cfg.block);
    const size_t shm = lds_bytes(cfg);
    hipStream_t s = as_stream(stream);
    const bool tiled = cfg.tile_width != 0;
#define RT_S(P, A, T)                                                                          \
    hipLaunchKernelGGL((k_trace_scene<P, A, T>), grid, block, shm, s, d_instances, count,      \
                       *rays, n, prim, t, d_inst, occl_bits, cfg.lds, tm)
    if (small_priv(depth, cfg.lds)) {
        if (any) { if (tiled) RT_S(kPrivSmall, true, true); else RT_S(kPrivSmall, true, false); }
        else if (tiled) RT_S(kPrivSmall, false, true);
        else RT_S(kPrivSmall, false, false);
    } else {
        if (any) { if (tiled) RT_S(kPrivLarge, true, true); else RT_S(kPrivLarge, true, false); }
        else if (tiled) RT_S(kPrivLarge, false, true);
        else RT_S(kPrivLarge, false, false);
    }
#undef RT_S
    return launch_status("k_trace_scene");
}

extern "C" rt_status rt_trace_scene(rt_stream stream, const rt_instance* d_instances,
                                    uint32_t count, uint32_t max_depth, const rt_rays_soa* rays,
                                    uint32_t n, uint32_t flags, rt_hits_soa* hits,
                                    uint32_t* d_inst, uint32_t* occl_bits) {
    return rt_trace_scene_ex(stream, d_instances, count, max_depth, rays, n, flags, nullptr,
                             hits, d_inst, occl_bits);
}
