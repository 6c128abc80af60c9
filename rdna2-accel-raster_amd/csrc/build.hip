// build.hip -- device BVH build on gfx950: centroid bounds -> 30-bit Morton
// codes -> stable LSD radix sort (8-bit digits, LDS-staged tiles, wave64
// ballot ranking) -> implicit complete Box4 tree in the fast_bvh_builder
// layout (include/radv_bvh.hpp:193-215: leaf boxes of 4 consecutive sorted
// triangles, parents of 4 consecutive boxes, root first) -> bottom-up AABB
// reduction.
//
// Two paths:
//  * n <= kSmallMax: one fused 1024-thread workgroup does everything in LDS
//    (the role of fast_bvh_builder.comp's single dispatch, with a sort).
//  * larger n: bounds, Morton, 4 x (histogram, digit scan, scatter) and
//    level-group emit kernels, each grid-wide.
#include "rt_internal.h"

namespace rt {

TreeLayout tree_layout(uint32_t n) {
    TreeLayout l;
    uint32_t c[32];
    uint32_t L = 0;
    c[L++] = (n + 3) / 4;
    while (c[L - 1] > 1) {
        c[L] = (c[L - 1] + 3) / 4;
        L++;
    }
    for (uint32_t i = 0; i < L; ++i) l.counts[i] = c[L - 1 - i];
    uint32_t off = 0;
    for (uint32_t i = 0; i < L; ++i) {
        l.offsets[i] = off;
        off += l.counts[i];
    }
    l.levels = L;
    l.total_boxes = off;
    l.total_as_size = 192ull + 128ull * off + 64ull * n;
    return l;
}

namespace {

constexpr uint32_t kSmallMax = 4096;   // fused single-workgroup path
constexpr int kSmallBlock = 1024;
constexpr int kSortBlock = 256;
constexpr int kSortItems = 16;
constexpr uint32_t kTile = kSortBlock * kSortItems;  // 4096 keys per tile
constexpr int kEmitBlock = 256;

struct LevelArgs {
    uint32_t counts[32];
    uint32_t offsets[32];
    uint32_t levels;
};

__device__ __forceinline__ void tri_verts(const float* __restrict__ pos,
                                          const uint32_t* __restrict__ idx, uint32_t t,
                                          float v[9]) {
    if (idx) {
        const uint32_t i0 = idx[3 * (size_t)t], i1 = idx[3 * (size_t)t + 1],
                       i2 = idx[3 * (size_t)t + 2];
        v[0] = pos[3 * (size_t)i0]; v[1] = pos[3 * (size_t)i0 + 1]; v[2] = pos[3 * (size_t)i0 + 2];
        v[3] = pos[3 * (size_t)i1]; v[4] = pos[3 * (size_t)i1 + 1]; v[5] = pos[3 * (size_t)i1 + 2];
        v[6] = pos[3 * (size_t)i2]; v[7] = pos[3 * (size_t)i2 + 1]; v[8] = pos[3 * (size_t)i2 + 2];
    } else {
        const float* p = pos + 9 * (size_t)t;
#pragma unroll
        for (int k = 0; k < 9; ++k) v[k] = p[k];
    }
}

// software_bvh.hpp:53 centroid: (p0 + p1 + p2) / 3
__device__ __forceinline__ void centroid(const float v[9], float c[3]) {
    c[0] = (v[0] + v[3] + v[6]) / 3.0f;
    c[1] = (v[1] + v[4] + v[7]) / 3.0f;
    c[2] = (v[2] + v[5] + v[8]) / 3.0f;
}

// q_a = clamp(floor((c_a - lo_a) * s_a), 0, 1023), s_a = 1024 / (hi_a - lo_a)
// (0 on a flat axis); x bits at 3k+2, y at 3k+1, z at 3k.
__device__ __forceinline__ uint32_t morton(const float c[3], const float lo[3],
                                           const float s[3]) {
    uint32_t q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float f = (c[a] - lo[a]) * s[a];
        f = fminf(fmaxf(f, 0.0f), 1023.0f);
        q[a] = (uint32_t)f;
    }
    return (expand_bits10(q[0]) << 2) | (expand_bits10(q[1]) << 1) | expand_bits10(q[2]);
}

__device__ __forceinline__ void axis_scale(const float lo[3], const float hi[3], float s[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float ext = hi[a] - lo[a];
        s[a] = ext > 0.0f ? 1024.0f / ext : 0.0f;
    }
}

// ---------------------------------------------------------------------------
// Block-wide primitives
// ---------------------------------------------------------------------------

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// Inclusive scan across the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t u = __shfl_up(v, o);
        if (lane >= (uint32_t)o) v += u;
    }
    return v;
}

// Stable rank of BLOCK*ITEMS elements by an 8-bit digit.  Element order is
// warp-striped: wave w, item i, lane l <-> position w*64*ITEMS + i*64 + l.
// On return pos[i] is the element's position in the digit-sorted tile and
// dstart[d] (LDS, 256 entries) the first position of digit d.  cnt is LDS of
// (BLOCK/64)*256 words, wsum LDS of 4 words.  Peer lanes sharing a digit are
// found with 8 ballots; each wave keeps its own running digit counters, so
// only the owning wave touches a counter row between the barriers.
template <int BLOCK, int ITEMS>
__device__ __forceinline__ void block_rank(const uint32_t (&dig)[ITEMS], uint32_t (&pos)[ITEMS],
                                           uint32_t* cnt, uint32_t* dstart, uint32_t* wsum) {
    constexpr int W = BLOCK / 64;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    for (uint32_t i = tid; i < (uint32_t)W * 256u; i += BLOCK) cnt[i] = 0;
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
    uint32_t* row = cnt + w * 256u;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t d = dig[i];
        uint64_t peers = ~0ull;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t below = __popcll(peers & lt);
        const uint32_t base = row[d];
        __builtin_amdgcn_wave_barrier();
        if (below == 0) row[d] = base + __popcll(peers);
        __builtin_amdgcn_wave_barrier();
        pos[i] = base + below;
    }
    __syncthreads();
    uint32_t total = 0;
    if (tid < 256) {
        uint32_t run = 0;
#pragma unroll
        for (int ww = 0; ww < W; ++ww) {
            const uint32_t c = cnt[ww * 256 + tid];
            cnt[ww * 256 + tid] = run;
            run += c;
        }
        total = run;
    }
    // exclusive scan of the 256 digit totals (threads 0..255 = 4 waves)
    uint32_t inc = 0;
    if (tid < 256) {
        inc = wave_incl_scan(total);
        if (lane == 63) wsum[w] = inc;
    }
    __syncthreads();
    if (tid < 256) {
        uint32_t add = 0;
        for (uint32_t ww = 0; ww < w; ++ww) add += wsum[ww];
        dstart[tid] = add + inc - total;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) pos[i] += dstart[dig[i]] + row[dig[i]];
}

// ---------------------------------------------------------------------------
// Large-n path
// ---------------------------------------------------------------------------

__global__ void k_init_bounds(uint32_t* ob) {
    if (threadIdx.x < 3) ob[threadIdx.x] = 0xFFFFFFFFu;      // min slots
    else if (threadIdx.x < 6) ob[threadIdx.x] = 0u;          // max slots
}

__global__ __launch_bounds__(256) void k_bounds(const float* __restrict__ pos,
                                                const uint32_t* __restrict__ idx, uint32_t n,
                                                uint32_t* __restrict__ ob) {
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t t = blockIdx.x * 256 + threadIdx.x; t < n; t += gridDim.x * 256) {
        float v[9], c[3];
        tri_verts(pos, idx, t, v);
        centroid(v, c);
#pragma unroll
        for (int a = 0; a < 3; ++a) { mn[a] = fminf(mn[a], c[a]); mx[a] = fmaxf(mx[a], c[a]); }
    }
    __shared__ float red[4][6];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float a0 = wave_min(mn[a]), a1 = wave_max(mx[a]);
        if (lane == 0) { red[w][a] = a0; red[w][3 + a] = a1; }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        float v = red[0][a];
        for (int ww = 1; ww < 4; ++ww) v = a < 3 ? fminf(v, red[ww][a]) : fmaxf(v, red[ww][a]);
        if (a < 3) atomicMin(ob + a, f2o(v));
        else atomicMax(ob + a, f2o(v));
    }
}

__global__ __launch_bounds__(256) void k_morton(const float* __restrict__ pos,
                                                const uint32_t* __restrict__ idx, uint32_t n,
                                                const uint32_t* __restrict__ ob,
                                                uint32_t* __restrict__ keys,
                                                uint32_t* __restrict__ vals) {
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    float lo[3] = {o2f(ob[0]), o2f(ob[1]), o2f(ob[2])};
    float hi[3] = {o2f(ob[3]), o2f(ob[4]), o2f(ob[5])};
    float s[3];
    axis_scale(lo, hi, s);
    float v[9], c[3];
    tri_verts(pos, idx, t, v);
    centroid(v, c);
    keys[t] = morton(c, lo, s);
    vals[t] = t;
}

// Per-tile digit histogram, digit-major: hist[d * tiles + tile].
__global__ __launch_bounds__(kSortBlock) void k_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                     uint32_t shift, uint32_t tiles,
                                                     uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kTile;
#pragma unroll 4
    for (int i = 0; i < kSortItems; ++i) {
        const uint32_t p = base + i * kSortBlock + threadIdx.x;
        if (p < n) atomicAdd(&h[(keys[p] >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// One block per digit: exclusive scan of that digit's tile counts in place,
// digit total to totals[d].
__global__ __launch_bounds__(1024) void k_digit_scan(uint32_t* __restrict__ hist, uint32_t tiles,
                                                     uint32_t* __restrict__ totals) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry;
    const uint32_t d = blockIdx.x, tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    uint32_t* row = hist + (size_t)d * tiles;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (uint32_t b = 0; b < tiles; b += 1024) {
        const uint32_t i = b + tid;
        const uint32_t v = i < tiles ? row[i] : 0u;
        const uint32_t inc = wave_incl_scan(v);
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        uint32_t add = carry;
        for (uint32_t ww = 0; ww < w; ++ww) add += wsum[ww];
        if (i < tiles) row[i] = add + inc - v;
        __syncthreads();
        if (tid == 1023) carry = add + inc;
        __syncthreads();
    }
    if (tid == 0) totals[d] = carry;
}

__global__ __launch_bounds__(kSortBlock) void k_scatter(const uint32_t* __restrict__ kin,
                                                        const uint32_t* __restrict__ vin,
                                                        uint32_t* __restrict__ kout,
                                                        uint32_t* __restrict__ vout, uint32_t n,
                                                        uint32_t shift, uint32_t tiles,
                                                        const uint32_t* __restrict__ hist,
                                                        const uint32_t* __restrict__ totals) {
    __shared__ uint32_t cnt[(kSortBlock / 64) * 256];
    __shared__ uint32_t dstart[256];
    __shared__ uint32_t gbase[256];
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t sk[kTile];
    __shared__ uint32_t sv[kTile];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    const uint32_t base = blockIdx.x * kTile;
    const uint32_t valid = min(kTile, n - base);

    // global base of each digit for this tile = exclusive(totals)[d] + hist[d][tile]
    {
        const uint32_t v = totals[tid];
        const uint32_t inc = wave_incl_scan(v);
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        uint32_t add = 0;
        for (uint32_t ww = 0; ww < w; ++ww) add += wsum[ww];
        gbase[tid] = add + inc - v + hist[(size_t)tid * tiles + blockIdx.x];
        __syncthreads();
    }
    uint32_t key[kSortItems], val[kSortItems], dig[kSortItems], pos[kSortItems];
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        const uint32_t lp = w * 64 * kSortItems + i * 64 + lane;
        const uint32_t p = base + lp;
        if (lp < valid) {
            key[i] = kin[p];
            val[i] = vin[p];
        } else {
            key[i] = 0xFFFFFFFFu;
            val[i] = 0u;
        }
        dig[i] = (key[i] >> shift) & 0xFFu;
    }
    block_rank<kSortBlock, kSortItems>(dig, pos, cnt, dstart, wsum);
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        sk[pos[i]] = key[i];
        sv[pos[i]] = val[i];
    }
    __syncthreads();
#pragma unroll 4
    for (int i = 0; i < kSortItems; ++i) {
        const uint32_t p = i * kSortBlock + tid;
        if (p < valid) {
            const uint32_t k = sk[p];
            const uint32_t d = (k >> shift) & 0xFFu;
            const uint32_t g = gbase[d] + (p - dstart[d]);
            kout[g] = k;
            vout[g] = sv[p];
        }
    }
}

__device__ __forceinline__ void write_slot(rt_node4* nd, int c, const float bb[6], int32_t child,
                                           int32_t count) {
    nd->minX[c] = bb[0]; nd->minY[c] = bb[1]; nd->minZ[c] = bb[2];
    nd->maxX[c] = bb[3]; nd->maxY[c] = bb[4]; nd->maxZ[c] = bb[5];
    nd->child[c] = child;
    nd->count[c] = count;
}

// Build one node in registers and store it as 8 x 16 B.
struct NodeReg {
    float4 mnx, mxx, mny, mxy, mnz, mxz;
    int4 ch, cn;
};
__device__ __forceinline__ void node_clear(NodeReg& r) {
    r.mnx = r.mxx = r.mny = r.mxy = r.mnz = r.mxz = make_float4(0.f, 0.f, 0.f, 0.f);
    r.ch = r.cn = make_int4(0, 0, 0, 0);
}
__device__ __forceinline__ void node_set(NodeReg& r, int c, const float bb[6], int32_t child,
                                         int32_t count) {
    float* a = reinterpret_cast<float*>(&r);
    a[0 + c] = bb[0]; a[4 + c] = bb[3];
    a[8 + c] = bb[1]; a[12 + c] = bb[4];
    a[16 + c] = bb[2]; a[20 + c] = bb[5];
    int32_t* ii = reinterpret_cast<int32_t*>(&r);
    ii[24 + c] = child;
    ii[28 + c] = count;
}
__device__ __forceinline__ void node_store(rt_node4* dst, const NodeReg& r) {
    float4* d = reinterpret_cast<float4*>(dst);
    d[0] = r.mnx; d[1] = r.mxx; d[2] = r.mny; d[3] = r.mxy; d[4] = r.mnz; d[5] = r.mxz;
    reinterpret_cast<int4*>(dst)[6] = r.ch;
    reinterpret_cast<int4*>(dst)[7] = r.cn;
}

// Leaf box b: sorted triangles 4b..4b+3 -> four rt_tri64 records and the
// leaf node (one triangle per slot, child = ~sorted position, count 1).
// REFIT takes each triangle's index from the existing leaf record.
template <bool REFIT>
__device__ __forceinline__ void emit_leaf_box(const float* __restrict__ pos,
                                              const uint32_t* __restrict__ idx, uint32_t n,
                                              const uint32_t* __restrict__ sorted,
                                              rt_node4* __restrict__ nodes,
                                              rt_tri64* __restrict__ tris, uint32_t leaf_off,
                                              uint32_t b, float box[6]) {
    NodeReg nr;
    node_clear(nr);
    box[0] = box[1] = box[2] = INFINITY;
    box[3] = box[4] = box[5] = -INFINITY;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const uint32_t s = 4 * b + c;
        if (s >= n) break;
        const uint32_t id = REFIT ? tris[s].triangle_id : sorted[s];
        float v[9];
        tri_verts(pos, idx, id, v);
        float4* tp = reinterpret_cast<float4*>(tris + s);
        tp[0] = make_float4(v[0], v[1], v[2], v[3]);
        tp[1] = make_float4(v[4], v[5], v[6], v[7]);
        tp[2] = make_float4(v[8], 0.f, 0.f, 0.f);
        reinterpret_cast<uint4*>(tp)[3] = make_uint4(id, 0x80000000u, 0u, 9u);
        float bb[6];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            // + 0.0f maps -0 to +0: which zero min/max return for (+0, -0)
            // differs between devices; canonical zeros keep the bytes exact.
            bb[a] = fminf(v[a], fminf(v[3 + a], v[6 + a])) + 0.0f;
            bb[3 + a] = fmaxf(v[a], fmaxf(v[3 + a], v[6 + a])) + 0.0f;
            box[a] = fminf(box[a], bb[a]);
            box[3 + a] = fmaxf(box[3 + a], bb[3 + a]);
        }
        node_set(nr, c, bb, (int32_t)~s, 1);
    }
    node_store(nodes + leaf_off + b, nr);
}

// Inner node i of level lv from child boxes (level lv+1, children 4i..4i+3).
__device__ __forceinline__ void emit_inner(rt_node4* __restrict__ nodes, const LevelArgs& L,
                                           uint32_t lv, uint32_t i, const float* child_boxes,
                                           uint32_t child_stride_base, float box[6]) {
    NodeReg nr;
    node_clear(nr);
    box[0] = box[1] = box[2] = INFINITY;
    box[3] = box[4] = box[5] = -INFINITY;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const uint32_t ch = 4 * i + c;
        if (ch >= L.counts[lv + 1]) break;
        const float* bb = child_boxes + 6 * (size_t)(ch - child_stride_base);
        float cb[6];
#pragma unroll
        for (int a = 0; a < 6; ++a) cb[a] = bb[a];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            box[a] = fminf(box[a], cb[a]);
            box[3 + a] = fmaxf(box[3 + a], cb[3 + a]);
        }
        node_set(nr, c, cb, (int32_t)(L.offsets[lv + 1] + ch), -1);
    }
    node_store(nodes + L.offsets[lv] + i, nr);
}

// Level group: bottom level lb (leaves when LEAF, else children boxes read
// from gboxes of level lb+1) plus up to 4 parent levels reduced in LDS.  Block
// b owns nodes [256b, 256b+256) of lb, [64b, 64b+64) of lb-1, ... .  The top
// level of the group writes its boxes to gboxes for the next launch.
template <bool LEAF, bool REFIT>
__global__ __launch_bounds__(kEmitBlock) void k_emit_levels(
    const float* __restrict__ pos, const uint32_t* __restrict__ idx, uint32_t n,
    const uint32_t* __restrict__ sorted, rt_node4* __restrict__ nodes,
    rt_tri64* __restrict__ tris, float* __restrict__ gboxes, LevelArgs L, uint32_t lb,
    uint32_t top) {
    __shared__ float sbox[2][kEmitBlock * 6];
    const uint32_t tid = threadIdx.x;
    uint32_t i = blockIdx.x * kEmitBlock + tid;
    float box[6];
    bool have = i < L.counts[lb];
    if (have) {
        if (LEAF) {
            emit_leaf_box<REFIT>(pos, idx, n, sorted, nodes, tris, L.offsets[lb], i, box);
        } else {
            emit_inner(nodes, L, lb, i, gboxes + 6 * (size_t)L.offsets[lb + 1], 0, box);
        }
    }
    if (lb == top) {
        if (have)
            for (int a = 0; a < 6; ++a) gboxes[6 * (size_t)(L.offsets[lb] + i) + a] = box[a];
        return;
    }
    int cur = 0;
    if (have)
        for (int a = 0; a < 6; ++a) sbox[cur][tid * 6 + a] = box[a];
    uint32_t width = kEmitBlock;
    for (uint32_t lv = lb; lv > top; --lv) {
        __syncthreads();
        width >>= 2;
        const uint32_t p = blockIdx.x * width + tid;  // node of level lv-1
        const bool act = tid < width && p < L.counts[lv - 1];
        if (act) {
            // children 4p..4p+3 of level lv live at LDS slots 4*tid + c
            emit_inner(nodes, L, lv - 1, p, sbox[cur], 4 * p - 4 * tid, box);
            if (lv - 1 == top) {
                for (int a = 0; a < 6; ++a) gboxes[6 * (size_t)(L.offsets[lv - 1] + p) + a] = box[a];
            } else {
                for (int a = 0; a < 6; ++a) sbox[cur ^ 1][tid * 6 + a] = box[a];
            }
        }
        cur ^= 1;
    }
}

// ---------------------------------------------------------------------------
// Small-n fused path: one 1024-thread workgroup, everything in LDS.
// ---------------------------------------------------------------------------
constexpr int kSmallItems = kSmallMax / kSmallBlock;  // 4
constexpr uint32_t kSmallBoxes = 1365;                // total boxes for n = 4096

struct SmallLds {
    uint32_t key[2][kSmallMax];
    uint32_t val[2][kSmallMax];
    uint32_t cnt[(kSmallBlock / 64) * 256];
    uint32_t dstart[256];
    uint32_t wsum[16];
    float red[16][6];
    float box[kSmallBoxes * 6];
};

template <bool REFIT>
__global__ __launch_bounds__(kSmallBlock) void k_build_small(
    const float* __restrict__ pos, const uint32_t* __restrict__ idx, uint32_t n,
    rt_node4* __restrict__ nodes, rt_tri64* __restrict__ tris, uint32_t* __restrict__ ids_out,
    uint32_t* __restrict__ codes_out, LevelArgs L) {
    __shared__ SmallLds S;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    int cur = 0;
    if (!REFIT) {
        // centroids of triangles p = w*256 + i*64 + lane (warp-striped order)
        float cen[kSmallItems][3];
        float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
        for (int i = 0; i < kSmallItems; ++i) {
            const uint32_t p = w * 64 * kSmallItems + i * 64 + lane;
            if (p < n) {
                float v[9];
                tri_verts(pos, idx, p, v);
                centroid(v, cen[i]);
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    mn[a] = fminf(mn[a], cen[i][a]);
                    mx[a] = fmaxf(mx[a], cen[i][a]);
                }
            }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float a0 = wave_min(mn[a]), a1 = wave_max(mx[a]);
            if (lane == 0) { S.red[w][a] = a0; S.red[w][3 + a] = a1; }
        }
        __syncthreads();
        float lo[3], hi[3], s[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = S.red[0][a];
            hi[a] = S.red[0][3 + a];
            for (int ww = 1; ww < 16; ++ww) {
                lo[a] = fminf(lo[a], S.red[ww][a]);
                hi[a] = fmaxf(hi[a], S.red[ww][3 + a]);
            }
        }
        axis_scale(lo, hi, s);
        uint32_t key[kSmallItems], val[kSmallItems], dig[kSmallItems], rk[kSmallItems];
#pragma unroll
        for (int i = 0; i < kSmallItems; ++i) {
            const uint32_t p = w * 64 * kSmallItems + i * 64 + lane;
            key[i] = p < n ? morton(cen[i], lo, s) : 0xFFFFFFFFu;
            val[i] = p;
        }
        for (uint32_t shift = 0; shift < 32; shift += 8) {
            if (shift > 0) {
#pragma unroll
                for (int i = 0; i < kSmallItems; ++i) {
                    const uint32_t p = w * 64 * kSmallItems + i * 64 + lane;
                    key[i] = S.key[cur][p];
                    val[i] = S.val[cur][p];
                }
                cur ^= 1;
            }
#pragma unroll
            for (int i = 0; i < kSmallItems; ++i) dig[i] = (key[i] >> shift) & 0xFFu;
            block_rank<kSmallBlock, kSmallItems>(dig, rk, S.cnt, S.dstart, S.wsum);
            const int dst = shift == 0 ? 0 : cur;
#pragma unroll
            for (int i = 0; i < kSmallItems; ++i) {
                S.key[dst][rk[i]] = key[i];
                S.val[dst][rk[i]] = val[i];
            }
            cur = dst;
            __syncthreads();
        }
        if (ids_out || codes_out) {
            for (uint32_t p = tid; p < n; p += kSmallBlock) {
                if (ids_out) ids_out[p] = S.val[cur][p];
                if (codes_out) codes_out[p] = S.key[cur][p];
            }
        }
    }
    // leaves + bottom-up levels, all boxes in LDS
    const uint32_t lb = L.levels - 1;
    if (tid < L.counts[lb]) {
        float box[6];
        emit_leaf_box<REFIT>(pos, idx, n, REFIT ? nullptr : S.val[cur], nodes, tris,
                             L.offsets[lb], tid, box);
        for (int a = 0; a < 6; ++a) S.box[6 * (L.offsets[lb] + tid) + a] = box[a];
    }
    for (int lv = (int)lb - 1; lv >= 0; --lv) {
        __syncthreads();
        if (tid < L.counts[lv]) {
            float box[6];
            emit_inner(nodes, L, lv, tid, S.box + 6 * L.offsets[lv + 1], 0, box);
            for (int a = 0; a < 6; ++a) S.box[6 * (L.offsets[lv] + tid) + a] = box[a];
        }
    }
}

LevelArgs level_args(const TreeLayout& t) {
    LevelArgs a;
    for (int i = 0; i < 32; ++i) {
        a.counts[i] = t.counts[i];
        a.offsets[i] = t.offsets[i];
    }
    a.levels = t.levels;
    return a;
}

struct ScratchPlan {
    size_t bounds = 0, keys0 = 0, vals0 = 0, keys1 = 0, vals1 = 0, hist = 0, totals = 0,
           boxes = 0, total = 0;
    uint32_t tiles = 0;
};

ScratchPlan plan_scratch(uint32_t n) {
    ScratchPlan p;
    const TreeLayout t = tree_layout(n);
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    size_t off = 0;
    p.tiles = ceil_div(n, kTile);
    p.bounds = off; off += al(8 * 4);
    p.keys0 = off; off += al((size_t)n * 4);
    p.vals0 = off; off += al((size_t)n * 4);
    p.keys1 = off; off += al((size_t)n * 4);
    p.vals1 = off; off += al((size_t)n * 4);
    p.hist = off; off += al((size_t)256 * p.tiles * 4);
    p.totals = off; off += al(256 * 4);
    p.boxes = off; off += al((size_t)t.total_boxes * 24);
    p.total = off;
    return p;
}

rt_status emit_levels(hipStream_t s, const float* pos, const uint32_t* idx, uint32_t n,
                      const uint32_t* sorted, rt_node4* nodes, rt_tri64* tris, float* gboxes,
                      const TreeLayout& t, bool refit) {
    const LevelArgs la = level_args(t);
    int lb = (int)t.levels - 1;
    bool leaf = true;
    while (lb >= 0) {
        const uint32_t top = (uint32_t)(lb >= 4 ? lb - 4 : 0);
        dim3 grid(ceil_div(t.counts[lb], kEmitBlock));
        if (leaf) {
            if (refit)
                hipLaunchKernelGGL((k_emit_levels<true, true>), grid, dim3(kEmitBlock), 0, s, pos,
                                   idx, n, sorted, nodes, tris, gboxes, la, (uint32_t)lb, top);
            else
                hipLaunchKernelGGL((k_emit_levels<true, false>), grid, dim3(kEmitBlock), 0, s,
                                   pos, idx, n, sorted, nodes, tris, gboxes, la, (uint32_t)lb,
                                   top);
        } else {
            hipLaunchKernelGGL((k_emit_levels<false, false>), grid, dim3(kEmitBlock), 0, s, pos,
                               idx, n, sorted, nodes, tris, gboxes, la, (uint32_t)lb, top);
        }
        rt_status st = launch_status("k_emit_levels");
        if (st != RT_OK) return st;
        leaf = false;
        lb = (int)top - 1;
    }
    return RT_OK;
}

}  // namespace
}  // namespace rt

using namespace rt;

extern "C" rt_status rt_tree_layout(uint32_t n, uint32_t* counts, uint32_t* offsets,
                                    uint32_t* levels, uint32_t* total_boxes,
                                    uint64_t* total_as_size) {
    const TreeLayout t = tree_layout(n);
    for (uint32_t i = 0; i < t.levels; ++i) {
        if (counts) counts[i] = t.counts[i];
        if (offsets) offsets[i] = t.offsets[i];
    }
    if (levels) *levels = t.levels;
    if (total_boxes) *total_boxes = t.total_boxes;
    if (total_as_size) *total_as_size = t.total_as_size;
    return RT_OK;
}

extern "C" rt_status rt_build_scratch_size(uint32_t n, size_t* scratch_bytes) {
    if (!scratch_bytes) return RT_ERR_ARG;
    *scratch_bytes = plan_scratch(n).total;
    return RT_OK;
}

extern "C" rt_status rt_gpu_build(rt_stream stream, const float* d_pos, const uint32_t* d_idx,
                                  uint32_t n, rt_node4* d_nodes, uint32_t node_capacity,
                                  rt_tri64* d_tris, uint32_t* d_sorted_ids, uint32_t* d_codes,
                                  void* d_scratch, size_t scratch_bytes) {
    if (n == 0) return RT_OK;
    if (!d_pos || !d_nodes || !d_tris) return RT_ERR_ARG;
    if (n > (1u << 30)) return RT_ERR_CAPACITY;
    const TreeLayout t = tree_layout(n);
    if (node_capacity < t.total_boxes) return RT_ERR_CAPACITY;
    hipStream_t s = as_stream(stream);
    if (n <= kSmallMax) {
        hipLaunchKernelGGL((k_build_small<false>), dim3(1), dim3(kSmallBlock), 0, s, d_pos, d_idx,
                           n, d_nodes, d_tris, d_sorted_ids, d_codes, level_args(t));
        return launch_status("k_build_small");
    }
    const ScratchPlan p = plan_scratch(n);
    if (!d_scratch || scratch_bytes < p.total) return RT_ERR_CAPACITY;
    char* base = static_cast<char*>(d_scratch);
    uint32_t* ob = reinterpret_cast<uint32_t*>(base + p.bounds);
    uint32_t* k0 = reinterpret_cast<uint32_t*>(base + p.keys0);
    uint32_t* v0 = reinterpret_cast<uint32_t*>(base + p.vals0);
    uint32_t* k1 = reinterpret_cast<uint32_t*>(base + p.keys1);
    uint32_t* v1 = reinterpret_cast<uint32_t*>(base + p.vals1);
    uint32_t* hist = reinterpret_cast<uint32_t*>(base + p.hist);
    uint32_t* totals = reinterpret_cast<uint32_t*>(base + p.totals);
    float* gboxes = reinterpret_cast<float*>(base + p.boxes);

    hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, s, ob);
    const uint32_t bgrid = std::min<uint32_t>(ceil_div(n, 256), 2048);
    hipLaunchKernelGGL(k_bounds, dim3(bgrid), dim3(256), 0, s, d_pos, d_idx, n, ob);
    hipLaunchKernelGGL(k_morton, dim3(ceil_div(n, 256)), dim3(256), 0, s, d_pos, d_idx, n, ob, k0,
                       v0);
    rt_status st = launch_status("k_morton");
    if (st != RT_OK) return st;
    uint32_t *ki = k0, *vi = v0, *ko = k1, *vo = v1;
    for (uint32_t shift = 0; shift < 32; shift += 8) {
        hipLaunchKernelGGL(k_hist, dim3(p.tiles), dim3(kSortBlock), 0, s, ki, n, shift, p.tiles,
                           hist);
        hipLaunchKernelGGL(k_digit_scan, dim3(256), dim3(1024), 0, s, hist, p.tiles, totals);
        hipLaunchKernelGGL(k_scatter, dim3(p.tiles), dim3(kSortBlock), 0, s, ki, vi, ko, vo, n,
                           shift, p.tiles, hist, totals);
        st = launch_status("k_scatter");
        if (st != RT_OK) return st;
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    // four passes: sorted data is back in k0 / v0
    if (d_sorted_ids) RT_TRY_HIP(hipMemcpyAsync(d_sorted_ids, v0, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
    if (d_codes) RT_TRY_HIP(hipMemcpyAsync(d_codes, k0, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
    return emit_levels(s, d_pos, d_idx, n, v0, d_nodes, d_tris, gboxes, t, false);
}

extern "C" rt_status rt_gpu_refit(rt_stream stream, const float* d_pos, const uint32_t* d_idx,
                                  uint32_t n, rt_node4* d_nodes, uint32_t node_count,
                                  rt_tri64* d_tris, void* d_scratch, size_t scratch_bytes) {
    if (n == 0) return RT_OK;
    if (!d_pos || !d_nodes || !d_tris) return RT_ERR_ARG;
    const TreeLayout t = tree_layout(n);
    if (node_count < t.total_boxes) return RT_ERR_CAPACITY;
    hipStream_t s = as_stream(stream);
    if (n <= kSmallMax) {
        hipLaunchKernelGGL((k_build_small<true>), dim3(1), dim3(kSmallBlock), 0, s, d_pos, d_idx,
                           n, d_nodes, d_tris, nullptr, nullptr, level_args(t));
        return launch_status("k_build_small");
    }
    const ScratchPlan p = plan_scratch(n);
    if (!d_scratch || scratch_bytes < p.total) return RT_ERR_CAPACITY;
    float* gboxes = reinterpret_cast<float*>(static_cast<char*>(d_scratch) + p.boxes);
    return emit_levels(s, d_pos, d_idx, n, nullptr, d_nodes, d_tris, gboxes, t, true);
}
