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
constexpr uint32_t kMaxDigits = 1024;  // widest radix digit: 10 bits

struct LevelArgs {
    uint32_t counts[32];
    uint32_t offsets[32];
    uint32_t levels;
};

__device__ __forceinline__ void tri_verts(const float* __restrict__ pos,
                                          const uint32_t* __restrict__ idx, uint32_t t,
                                          float v[9]) {
    // one pointer per vertex (both layouts), each read at offsets 0, 4, 8
    const float *p0, *p1, *p2;
    if (idx) {
        const uint32_t i0 = idx[3 * (size_t)t], i1 = idx[3 * (size_t)t + 1],
                       i2 = idx[3 * (size_t)t + 2];
        p0 = pos + 3 * (size_t)i0;
        p1 = pos + 3 * (size_t)i1;
        p2 = pos + 3 * (size_t)i2;
    } else {
        p0 = pos + 9 * (size_t)t;
        p1 = p0 + 3;
        p2 = p0 + 6;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        v[k] = p0[k];
        v[3 + k] = p1[k];
        v[6 + k] = p2[k];
    }
}

// software_bvh.hpp:53 centroid: (p0 + p1 + p2) / 3 (div3 == the division)
__device__ __forceinline__ void centroid(const float v[9], float c[3]) {
    c[0] = div3(v[0] + v[3] + v[6]);
    c[1] = div3(v[1] + v[4] + v[7]);
    c[2] = div3(v[2] + v[5] + v[8]);
}

// q_a = clamp(floor((c_a - lo_a) * s_a), 0, 1023), s_a = 1024 / (hi_a - lo_a)
// (0 on a flat axis); x bits at 3k+2, y at 3k+1, z at 3k.
__device__ __forceinline__ void quantize(const float c[3], const float lo[3], const float s[3],
                                         uint32_t q[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float f = (c[a] - lo[a]) * s[a];
        f = fminf(fmaxf(f, 0.0f), 1023.0f);
        q[a] = (uint32_t)f;
    }
}
__device__ __forceinline__ uint32_t morton(const float c[3], const float lo[3],
                                           const float s[3]) {
    uint32_t q[3];
    quantize(c, lo, s, q);
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

// Inclusive wave scan with DPP (the gfx9 sequence): row_shr 1, 2, 4, 8 scan
// each row of 16 lanes, row_bcast:15 and row_bcast:31 carry across rows.
// All 64 lanes must be active; lane i ends with op over lanes 0..i.  (A
// shuffle-based scan goes through ds_bpermute, an LDS round trip per step.)
template <typename Op>
__device__ __forceinline__ uint32_t wave_scan_dpp(uint32_t v, uint32_t id, Op op) {
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, 0x111, 0xf, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, 0x112, 0xf, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, 0x114, 0xf, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, 0x118, 0xf, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, 0x142, 0xa, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, 0x143, 0xc, 0xf, false));
    return v;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    return wave_scan_dpp(v, 0u, [](uint32_t a, uint32_t b) { return a + b; });
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    v = wave_scan_dpp(v, 0u, [](uint32_t a, uint32_t b) { return a > b ? a : b; });
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}
// Float min / max through the order-preserving integer map (so -0 < +0; the
// callers only quantize against the result, where the sign of zero is moot).
__device__ __forceinline__ float wave_min(float f) {
    uint32_t v = wave_scan_dpp(f2o(f), 0xFFFFFFFFu,
                               [](uint32_t a, uint32_t b) { return a < b ? a : b; });
    return o2f((uint32_t)__builtin_amdgcn_readlane((int)v, 63));
}
__device__ __forceinline__ float wave_max(float f) {
    uint32_t v = wave_scan_dpp(f2o(f), 0u, [](uint32_t a, uint32_t b) { return a > b ? a : b; });
    return o2f((uint32_t)__builtin_amdgcn_readlane((int)v, 63));
}

// Exclusive scan of D per-digit values held in LDS or registers by the
// block's first threads: thread t < ST owns the R = D / ST consecutive digits
// [t*R, t*R + R) (ST = min(D, 256)), passes their values in v[] and gets
// their exclusive prefixes back.  wsum: 4 words of LDS.  Every thread of the
// block calls it (it contains a barrier).
template <uint32_t D>
__device__ __forceinline__ void digit_excl_scan(uint32_t (&v)[(D > 256 ? D / 256 : 1)],
                                                uint32_t* wsum) {
    constexpr uint32_t R = D > 256 ? D / 256 : 1, ST = D / R, SW = (ST + 63) / 64;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    uint32_t local = 0;
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) local += tid < ST ? v[r] : 0u;
    uint32_t inc = 0;
    if (tid < SW * 64) {
        inc = wave_incl_scan(local);
        if (lane == 63) wsum[w] = inc;
    }
    lds_barrier();
    if (tid < ST) {
        uint32_t run = inc - local;
        for (uint32_t ww = 0; ww < w; ++ww) run += wsum[ww];
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) {
            const uint32_t x = v[r];
            v[r] = run;
            run += x;
        }
    }
}

// Stable rank of BLOCK*ITEMS elements by a BITS-bit digit.  Element order is
// warp-striped: wave w, item i, lane l <-> position w*64*ITEMS + i*64 + l.
// On return pos[i] is the element's position in the digit-sorted tile and
// dstart[d] (LDS, 2^BITS entries) the first position of digit d.  cnt is LDS
// of (BLOCK/64)*2^BITS words, wsum LDS of 4 words.  Peer lanes sharing a
// digit are found with BITS ballots; each wave keeps its own running digit
// counters, so only the owning wave touches a counter row between the
// barriers.
template <int BLOCK, int ITEMS, int BITS = 8>
__device__ __forceinline__ void block_rank(const uint32_t (&dig)[ITEMS], uint32_t (&pos)[ITEMS],
                                           uint32_t* cnt, uint32_t* dstart, uint32_t* wsum) {
    constexpr uint32_t D = 1u << BITS, W = BLOCK / 64;
    constexpr uint32_t R = D > 256 ? D / 256 : 1, ST = D / R;
    static_assert(ST <= (uint32_t)BLOCK, "one scanning thread per R digits");
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    for (uint32_t i = tid; i < W * D; i += BLOCK) cnt[i] = 0;
    lds_barrier();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
    uint32_t* row = cnt + w * D;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t d = dig[i];
        uint64_t peers = ~0ull;
#pragma unroll
        for (int b = 0; b < BITS; ++b) {
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
    lds_barrier();
    // per digit: exclusive prefix over the waves' rows, then over digits
    uint32_t tot[R];
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) {
        tot[r] = 0;
        if (tid < ST) {
            const uint32_t d = tid * R + r;
            uint32_t run = 0;
#pragma unroll
            for (uint32_t ww = 0; ww < W; ++ww) {
                const uint32_t c = cnt[ww * D + d];
                cnt[ww * D + d] = run;
                run += c;
            }
            tot[r] = run;
        }
    }
    digit_excl_scan<D>(tot, wsum);
    if (tid < ST) {
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) dstart[tid * R + r] = tot[r];
    }
    lds_barrier();
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

// Per-tile digit histogram of a BITS-bit digit, digit-major:
// hist[d * tiles + tile].
template <int BITS>
__global__ __launch_bounds__(kSortBlock) void k_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                     uint32_t shift, uint32_t tiles,
                                                     uint32_t* __restrict__ hist) {
    constexpr uint32_t D = 1u << BITS;
    __shared__ uint32_t h[D];
    for (uint32_t d = threadIdx.x; d < D; d += kSortBlock) h[d] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kTile;
#pragma unroll 4
    for (int i = 0; i < kSortItems; ++i) {
        const uint32_t p = base + i * kSortBlock + threadIdx.x;
        if (p < n) atomicAdd(&h[(keys[p] >> shift) & (D - 1u)], 1u);
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < D; d += kSortBlock) hist[d * tiles + blockIdx.x] = h[d];
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

template <int BITS>
__global__ __launch_bounds__(kSortBlock) void k_scatter(const uint32_t* __restrict__ kin,
                                                        const uint32_t* __restrict__ vin,
                                                        uint32_t* __restrict__ kout,
                                                        uint32_t* __restrict__ vout, uint32_t n,
                                                        uint32_t shift, uint32_t tiles,
                                                        const uint32_t* __restrict__ hist,
                                                        const uint32_t* __restrict__ totals) {
    constexpr uint32_t D = 1u << BITS;
    constexpr uint32_t R = D > 256 ? D / 256 : 1, ST = D / R;
    __shared__ uint32_t cnt[(kSortBlock / 64) * D];
    __shared__ uint32_t dstart[D];
    __shared__ uint32_t gbase[D];
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t sk[kTile];
    __shared__ uint32_t sv[kTile];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    const uint32_t base = blockIdx.x * kTile;
    const uint32_t valid = min(kTile, n - base);

    // global base of each digit for this tile = exclusive(totals)[d] + hist[d][tile]
    {
        uint32_t v[R];
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) v[r] = tid < ST ? totals[tid * R + r] : 0u;
        digit_excl_scan<D>(v, wsum);
        if (tid < ST) {
#pragma unroll
            for (uint32_t r = 0; r < R; ++r) {
                const uint32_t d = tid * R + r;
                gbase[d] = v[r] + hist[(size_t)d * tiles + blockIdx.x];
            }
        }
        lds_barrier();
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
        dig[i] = (key[i] >> shift) & (D - 1u);
    }
    block_rank<kSortBlock, kSortItems, BITS>(dig, pos, cnt, dstart, wsum);
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
            const uint32_t d = (k >> shift) & (D - 1u);
            const uint32_t g = gbase[d] + (p - dstart[d]);
            kout[g] = k;
            vout[g] = sv[p];
        }
    }
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

// 64-B leaf record (rt_tri64 / RadvBvhTriangleNode): vertices, triangle id,
// geometry flags 0x80000000, node id 9.
__device__ __forceinline__ void store_tri(rt_tri64* dst, const float v[9], uint32_t id) {
    float4* tp = reinterpret_cast<float4*>(dst);
    tp[0] = make_float4(v[0], v[1], v[2], v[3]);
    tp[1] = make_float4(v[4], v[5], v[6], v[7]);
    tp[2] = make_float4(v[8], 0.f, 0.f, 0.f);
    reinterpret_cast<uint4*>(tp)[3] = make_uint4(id, 0x80000000u, 0u, 9u);
}

// Triangle box.  + 0.0f maps -0 to +0: which zero min/max return for
// (+0, -0) differs between devices; canonical zeros keep the bytes exact.
__device__ __forceinline__ void tri_box(const float v[9], float bb[6]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        bb[a] = fminf(v[a], fminf(v[3 + a], v[6 + a])) + 0.0f;
        bb[3 + a] = fmaxf(v[a], fmaxf(v[3 + a], v[6 + a])) + 0.0f;
    }
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
        store_tri(tris + s, v, id);
        float bb[6];
        tri_box(v, bb);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            box[a] = fminf(box[a], bb[a]);
            box[3 + a] = fmaxf(box[3 + a], bb[3 + a]);
        }
        node_set(nr, c, bb, (int32_t)~s, 1);
    }
    node_store(nodes + leaf_off + b, nr);
}

// Inner node i of level lv from child boxes (level lv+1, children 4i..4i+3).
// LT: a level table with counts[] / offsets[] (LevelArgs, MedTree).
template <typename LT>
__device__ __forceinline__ void emit_inner(rt_node4* __restrict__ nodes, const LT& L,
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
                for (int a = 0; a < 6; ++a)
                    gboxes[6 * (size_t)(L.offsets[lv - 1] + p) + a] = box[a];
            } else {
                for (int a = 0; a < 6; ++a) sbox[cur ^ 1][tid * 6 + a] = box[a];
            }
        }
        cur ^= 1;
    }
}

// ---------------------------------------------------------------------------
// Small-n fused path: n <= 4096 in one launch of ceil(n / 256) workgroups of
// 1024 threads.  Every workgroup loads all triangles, sorts them and reduces
// every box in its own LDS (identical work, identical results); workgroup k
// then writes only its share of the output: the records, sorted indices and
// codes of sorted positions [256k, 256k + 256) and 1/ceil(n/256) of the node
// array.  One CU stores about 57 GB/s even fully coalesced, so the 463 KB of
// output of a 4096-triangle build is spread over 16 CUs rather than 1.
//
// Thread t owns triangles p = w*256 + i*64 + lane (w = t/64, i < 4) and keeps
// their 36 vertex floats in registers from the load to the output.  The
// centroid divide is div3 (the IEEE division sequence's temporaries pushed
// the load phase past 128 VGPRs and into spills).  The sort is one MSD
// bucket pass over the top 13 code bits (LDS histogram, scan, scatter); a
// triangle's rank inside its bucket is the number of members with a smaller
// (low 17 code bits, triangle index) pair, one u32 since n <= 4096.  That is
// the order (code, index), i.e. the order of a stable sort by code.  When a
// bucket holds more than kBucketCap triangles (clustered centroids) the
// workgroup runs the 4-pass LSD ballot sort instead.
//
// LDS boxes are SoA (one array per bound) with level offsets padded to 4, so
// a node's 4 child bounds are one 16-B read; stores put consecutive lanes on
// consecutive 16-B chunks.
// ---------------------------------------------------------------------------
constexpr int kSmallItems = kSmallMax / kSmallBlock;  // 4
constexpr uint32_t kBoxPitch = 1368;  // padded box slots for n = 4096: 4+4+16+64+256+1024
constexpr uint32_t kBucketBits = 13;
constexpr uint32_t kBuckets = 1u << kBucketBits;
constexpr uint32_t kBucketShift = 30 - kBucketBits;
constexpr uint32_t kBucketCap = 256;
constexpr uint32_t kPartRecs = 256;   // sorted positions written per workgroup

// Level table of a fused-path tree (n <= 4096: at most 6 levels).  Read only
// with compile-time indices (pick), so it stays in scalar registers; a
// runtime index into a kernel argument array is a scalar memory load.
constexpr int kSmallLevels = 7;
struct SmallTree {
    uint32_t levels;              // node levels, root first
    uint32_t count[kSmallLevels];  // nodes per level
    uint32_t off[kSmallLevels];    // node index of the level's first node
    uint32_t pad[kSmallLevels];    // first LDS box slot of the level (multiple of 4)
};

// The level table lives in LDS for the kernel's lifetime: lookups are
// wave-uniform broadcast reads.  (Kept in a private array it went to
// scratch: indexing a kernel-argument or register array by a runtime level
// compiles to memory loads either way.)
struct LevelLds {
    uint32_t levels;
    uint32_t count[kSmallLevels];
    uint32_t off[kSmallLevels];
    uint32_t pad[kSmallLevels];
};

// Threads 0..kSmallLevels-1 copy the table; visible after the next barrier.
__device__ __forceinline__ void load_levels(LevelLds& lv, const SmallTree& T) {
    const uint32_t q = threadIdx.x;
    if (q < (uint32_t)kSmallLevels) {
        lv.count[q] = T.count[q];
        lv.off[q] = T.off[q];
        lv.pad[q] = T.pad[q];
        if (q == 0) lv.levels = T.levels;
    }
}

struct SmallLds {
    union {
        struct {
            uint32_t start[kBuckets];   // bucket counts, then first ranks
            uint32_t cursor[kBuckets];  // scatter cursors, then bucket ends
            uint32_t comp[kSmallMax];   // (low 17 code bits << 12) | index
        } msd;
        struct {
            uint32_t key[2][kSmallMax];
            uint32_t val[2][kSmallMax];
            uint32_t cnt[(kSmallBlock / 64) * 256];
        } lsd;
        float tbox[6][kSmallMax];                  // triangle bounds at their ranks
    } u;
    float box[6][kBoxPitch];  // node bounds: level lv from slot pad[lv]
    uint32_t spread[1024];    // expand_bits10 of every 10-bit value
    LevelLds lv;              // level table of the mesh being built
    uint32_t dstart[256];
    uint32_t wsum[16];
    uint32_t wmax[16];
    float red[16][6];
};


// Union of slots first..first+3 of src (SoA, pitch floats per bound, first a
// multiple of 4).  Empty slots hold the identity box (+inf mins, -inf maxes),
// so no slot is masked.
__device__ __forceinline__ void union4(const float* src, uint32_t pitch, uint32_t first,
                                       float out[6]) {
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        const float4 q = *reinterpret_cast<const float4*>(src + a * pitch + first);
        out[a] = a < 3 ? fminf(fminf(q.x, q.y), fminf(q.z, q.w))
                       : fmaxf(fmaxf(q.x, q.y), fmaxf(q.z, q.w));
    }
}

__device__ __forceinline__ void put_box(float (*box)[kBoxPitch], uint32_t slot, const float b[6]) {
#pragma unroll
    for (int a = 0; a < 6; ++a) box[a][slot] = b[a];
}

// Node j of a level whose children are slots 4j..4j+3 of src, or the identity
// box for a padding slot j >= count.
__device__ __forceinline__ void level_box(const float* src, uint32_t pitch, uint32_t j,
                                          uint32_t count, float b[6]) {
    if (j < count) {
        union4(src, pitch, 4 * j, b);
    } else {
        b[0] = b[1] = b[2] = INFINITY;
        b[3] = b[4] = b[5] = -INFINITY;
    }
}

// Every node box, from the triangle boxes in tbox (identity-padded to a
// multiple of 4): leaves by all threads, the level above them (at most 256
// nodes) by the first 4 waves, the rest (at most 64 nodes per level) by wave
// 0 alone.  Each level is padded with identity boxes to a multiple of 4.
// Ends with a barrier.
__device__ void small_boxes(SmallLds& S) {
    const LevelLds& L = S.lv;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    const uint32_t lb = L.levels - 1;
    {
        const uint32_t c = L.count[lb];
        if (tid < ((c + 3u) & ~3u)) {
            float b[6];
            level_box(S.u.tbox[0], kSmallMax, tid, c, b);
            put_box(S.box, L.pad[lb] + tid, b);
        }
    }
    lds_barrier();
    if (lb == 0) return;
    {
        const uint32_t c = L.count[lb - 1];
        if (tid < ((c + 3u) & ~3u)) {
            float b[6];
            level_box(S.box[0] + L.pad[lb], kBoxPitch, tid, c, b);
            put_box(S.box, L.pad[lb - 1] + tid, b);
        }
    }
    lds_barrier();
    if (lb == 1) return;
    if (w == 0) {
        for (int lv = (int)lb - 2; lv >= 0; --lv) {
            wave_lds_sync();
            const uint32_t c = L.count[lv];
            if (lane < ((c + 3u) & ~3u)) {
                float b[6];
                level_box(S.box[0] + L.pad[lv + 1], kBoxPitch, lane, c, b);
                put_box(S.box, L.pad[lv] + lane, b);
            }
        }
    }
    lds_barrier();
}

// Chunk c (16 B) of node `node`: c < 6 one bound of the 4 slots (minX, maxX,
// minY, maxY, minZ, maxZ: software_bvh.hpp:178-187 order), 6 child indices,
// 7 child counts.  Leaf slot s = 4j + k: triangle bounds, child ~s, count 1;
// inner slot: child bounds, child node index, count -1; empty slots zero.
__device__ __forceinline__ uint4 node_chunk(uint32_t n, uint32_t node, uint32_t c,
                                            const SmallLds& S) {
    const LevelLds& L = S.lv;
    const uint32_t lb = L.levels - 1;
    uint32_t lv = 0;
#pragma unroll
    for (int q = 1; q < kSmallLevels; ++q)
        if ((uint32_t)q < L.levels && node >= L.off[q]) lv = q;
    const uint32_t j = node - L.off[lv];
    const bool leaf = lv == lb;
    const uint32_t cnt = leaf ? n : L.count[lv + 1];
    const uint32_t next = L.off[lv + 1];
    uint32_t o[4] = {0u, 0u, 0u, 0u};
    if (c < 6) {
        const uint32_t comp = (c & 1u) * 3u + (c >> 1);  // 0,3,1,4,2,5
        const float* src = leaf ? S.u.tbox[comp] + 4 * j : S.box[comp] + L.pad[lv + 1] + 4 * j;
        const float4 q = *reinterpret_cast<const float4*>(src);
        o[0] = __float_as_uint(q.x); o[1] = __float_as_uint(q.y);
        o[2] = __float_as_uint(q.z); o[3] = __float_as_uint(q.w);
    }
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        const uint32_t ch = 4 * j + k;
        if (ch >= cnt) o[k] = 0u;
        else if (c == 6) o[k] = leaf ? ~ch : next + ch;
        else if (c == 7) o[k] = leaf ? 1u : 0xFFFFFFFFu;
    }
    return make_uint4(o[0], o[1], o[2], o[3]);
}

// Chunk c (16 B) of box node `node` in the RADV image
// (radv_bvh.hpp:79-135, written by fast_bvh_builder.comp:141-212): chunk 0
// the 4 child node IDs (byte offset from the first box, >> 3, | type: 0
// triangle, 5 box32), chunks 1-6 the children's AABBs as 24-B records (min
// xyz, max xyz), chunk 7 zero.  Empty slot: ID 0xFFFFFFFF, NaN box.
__device__ __forceinline__ uint4 radv_chunk(uint32_t n, uint32_t node, uint32_t c,
                                            const SmallLds& S) {
    const LevelLds& L = S.lv;
    const uint32_t lb = L.levels - 1;
    uint32_t lv = 0;
#pragma unroll
    for (int q = 1; q < kSmallLevels; ++q)
        if ((uint32_t)q < L.levels && node >= L.off[q]) lv = q;
    const uint32_t j = node - L.off[lv];
    const bool leaf = lv == lb;
    const uint32_t cnt = leaf ? n : L.count[lv + 1];
    const uint32_t tb = L.off[lb] + L.count[lb];
    uint32_t o[4] = {0u, 0u, 0u, 0u};
    if (c == 0) {
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            const uint32_t ch = 4 * j + k;
            o[k] = ch >= cnt ? 0xFFFFFFFFu
                 : leaf      ? (128u * tb + 64u * ch) >> 3
                             : (((L.off[lv + 1] + ch) * 128u) >> 3) | 5u;
        }
    } else if (c < 7) {
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) {
            const uint32_t e = 4 * (c - 1) + q, k = e / 6, a = e % 6, ch = 4 * j + k;
            const float* src = leaf ? S.u.tbox[a] : S.box[a] + L.pad[lv + 1];
            o[q] = ch < cnt ? __float_as_uint(src[ch]) : 0x7FC00000u;
        }
    }
    return make_uint4(o[0], o[1], o[2], o[3]);
}

// The 192-B RADV header and pre-root links (fast_bvh_builder.comp:220-256):
// root bounds, sizes, counts and the pre-root box's links at 0xB4 / 0xBC.
__device__ void radv_header(uint32_t* __restrict__ h, uint32_t n, const SmallLds& S) {
    const LevelLds& L = S.lv;
    const uint32_t lb = L.levels - 1, tb = L.off[lb] + L.count[lb];
    const uint64_t sz = 192ull + 128ull * tb + 64ull * n;
    uint32_t w[48] = {0u};
    w[0] = 192u;
    for (int a = 0; a < 3; ++a) {
        w[2 + a] = __float_as_uint(S.box[a][L.pad[0]]);
        w[5 + a] = __float_as_uint(S.box[3 + a][L.pad[0]]);
    }
    w[8] = 8u; w[9] = 1u; w[10] = 1u;
    w[12] = (uint32_t)sz; w[13] = (uint32_t)(sz >> 32);
    w[14] = (uint32_t)(sz + 128); w[15] = (uint32_t)((sz + 128) >> 32);
    w[16] = 1u; w[17] = 1u; w[18] = 1u;
    w[20] = (uint32_t)sz; w[21] = (uint32_t)(sz >> 32);
    w[23] = 1u;
    const uint64_t tri0 = 192ull + 128ull * tb;
    w[24] = (uint32_t)tri0; w[25] = (uint32_t)(tri0 >> 32);
    w[26] = n;
    w[31] = 4u;
    w[32] = 1u;
    w[45] = 5u;
    w[47] = 0xFFFFFFFFu;
    uint4* d = reinterpret_cast<uint4*>(h);
#pragma unroll
    for (int k = 0; k < 12; ++k)
        d[k] = make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]);
}

// Output of the fused build and refit.  Owner item i (ok[i]): vertices v[i],
// sorted position r[i], triangle index id[i], code key[i].  Every workgroup
// computes all boxes; workgroup `part` of `parts` stores its share.
template <bool RADV>
__device__ void small_emit(SmallLds& S, uint32_t n,
                           const float (&v)[kSmallItems][9], const uint32_t (&r)[kSmallItems],
                           const uint32_t (&id)[kSmallItems], const uint32_t (&key)[kSmallItems],
                           const bool (&ok)[kSmallItems], rt_node4* __restrict__ nodes,
                           rt_tri64* __restrict__ tris, uint32_t* __restrict__ ids_out,
                           uint32_t* __restrict__ codes_out, uint32_t part, uint32_t parts) {
    const uint32_t tid = threadIdx.x;
    const LevelLds& L = S.lv;
#pragma unroll
    for (int i = 0; i < kSmallItems; ++i) {
        if (!ok[i]) continue;
        float bb[6];
        tri_box(v[i], bb);
#pragma unroll
        for (int a = 0; a < 6; ++a) S.u.tbox[a][r[i]] = bb[a];
    }
    if (tid < ((n + 3u) & ~3u) - n) {  // identity boxes up to a multiple of 4
#pragma unroll
        for (int a = 0; a < 6; ++a) S.u.tbox[a][n + tid] = a < 3 ? INFINITY : -INFINITY;
    }
    lds_barrier();
    small_boxes(S);
    const uint32_t lb = L.levels - 1;
    // this workgroup's slice of the node array, 16-B chunks
    const uint32_t nchunks = (L.off[lb] + L.count[lb]) * 8;
    const uint32_t per = (nchunks + parts - 1) / parts;
    const uint32_t g0 = part * per, g1 = min(nchunks, g0 + per);
    uint4* nd = reinterpret_cast<uint4*>(nodes);
    for (uint32_t g = g0 + tid; g < g1; g += kSmallBlock)
        nd[g] = RADV ? radv_chunk(n, g >> 3, g & 7u, S) : node_chunk(n, g >> 3, g & 7u, S);
    if (RADV && part == 0 && tid == 0)  // the image starts 192 B before the first box
        radv_header(reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(nodes) - 192), n, S);
    // records: the thread holding a triangle writes its 64-B record (4 x 16 B,
    // store_tri's layout) when its sorted position lies in this workgroup's
    // share: slices of kPartRecs positions, every parts-th one (all of them
    // when one workgroup builds the mesh alone)
#pragma unroll
    for (int i = 0; i < kSmallItems; ++i) {
        if (!ok[i] || (r[i] / kPartRecs) % parts != part) continue;
        uint4* td = reinterpret_cast<uint4*>(tris + r[i]);
        uint32_t u[9];
#pragma unroll
        for (int a = 0; a < 9; ++a) u[a] = __float_as_uint(v[i][a]);
        td[0] = make_uint4(u[0], u[1], u[2], u[3]);
        td[1] = make_uint4(u[4], u[5], u[6], u[7]);
        td[2] = make_uint4(u[8], 0u, 0u, 0u);
        td[3] = make_uint4(id[i], 0x80000000u, 0u, 9u);
        if (ids_out) ids_out[r[i]] = id[i];
        if (codes_out) codes_out[r[i]] = key[i];
    }
}

__device__ __forceinline__ void small_info(rt_build_info* info, const SmallTree& T,
                                           uint32_t part) {
    if (info && part == 0 && threadIdx.x == 0) {
        info->node_count = T.off[T.levels - 1] + T.count[T.levels - 1];
        info->depth = T.levels;
    }
}

// The whole fused build of one mesh by this workgroup, which writes slice
// `part` of `parts` of the output (see small_emit).
template <bool RADV>
__device__ void build_small(SmallLds& S, const SmallTree& T, const float* __restrict__ pos,
                            const uint32_t* __restrict__ idx, uint32_t n,
                            rt_node4* __restrict__ nodes, rt_tri64* __restrict__ tris,
                            uint32_t* __restrict__ ids_out, uint32_t* __restrict__ codes_out,
                            rt_build_info* info, uint32_t part, uint32_t parts,
                            bool sort = true) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    small_info(info, T, part);
    load_levels(S.lv, T);  // first read after the barriers below
    // bucket counters cleared while the triangle loads are in flight
    static_assert(kBuckets == 8 * kSmallBlock, "8 bucket counters per thread");
    uint4* start4 = reinterpret_cast<uint4*>(S.u.msd.start);
    start4[2 * tid] = make_uint4(0u, 0u, 0u, 0u);
    start4[2 * tid + 1] = make_uint4(0u, 0u, 0u, 0u);
    S.spread[tid] = expand_bits10(tid);

    float v[kSmallItems][9], cen[kSmallItems][3];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
    for (int i = 0; i < kSmallItems; ++i) {
        const uint32_t p = w * 64 * kSmallItems + i * 64 + lane;
        if (p < n) {
            tri_verts(pos, idx, p, v[i]);
            centroid(v[i], cen[i]);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                mn[a] = fminf(mn[a], cen[i][a]);
                mx[a] = fmaxf(mx[a], cen[i][a]);
            }
        }
    }
    uint32_t key[kSmallItems], rank[kSmallItems];
    if (!sort) {  // input order (fast_bvh_builder's leaves of consecutive triangles)
#pragma unroll
        for (int i = 0; i < kSmallItems; ++i) {
            key[i] = 0u;
            rank[i] = w * 64 * kSmallItems + i * 64 + lane;
        }
    } else {
        // bounds: the 1024 per-thread partials through LDS (the box arrays are
        // free until the output), bound a reduced by wave a alone
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            S.box[a][tid] = mn[a];
            S.box[3 + a][tid] = mx[a];
        }
        lds_barrier();
        if (w < 6) {
            const float4* src = reinterpret_cast<const float4*>(S.box[w] + 16 * lane);
            float m = w < 3 ? INFINITY : -INFINITY;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float4 x = src[q];
                m = w < 3 ? fminf(m, fminf(fminf(x.x, x.y), fminf(x.z, x.w)))
                          : fmaxf(m, fmaxf(fmaxf(x.x, x.y), fmaxf(x.z, x.w)));
            }
            m = w < 3 ? wave_min(m) : wave_max(m);
            if (lane == 0) S.red[0][w] = m;
        }
        lds_barrier();
        float lo[3], hi[3], sc[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = S.red[0][a];
            hi[a] = S.red[0][3 + a];
        }
        axis_scale(lo, hi, sc);
#pragma unroll
        for (int i = 0; i < kSmallItems; ++i) {
            const uint32_t p = w * 64 * kSmallItems + i * 64 + lane;
            uint32_t q[3];
            quantize(cen[i], lo, sc, q);
            key[i] = p < n ? (S.spread[q[0]] << 2) | (S.spread[q[1]] << 1) | S.spread[q[2]]
                           : 0xFFFFFFFFu;
            if (p < n) atomicAdd(&S.u.msd.start[key[i] >> kBucketShift], 1u);
        }
        lds_barrier();
        // exclusive scan of the bucket counts (8 per thread) and the largest bucket
        {
            const uint4 c0 = start4[2 * tid], c1 = start4[2 * tid + 1];
            const uint32_t c[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
            uint32_t total = 0, big = 0;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                total += c[k];
                big = max(big, c[k]);
            }
            const uint32_t inc = wave_incl_scan(total);
            big = wave_max_u32(big);
            if (lane == 63) S.wsum[w] = inc;
            if (lane == 0) S.wmax[w] = big;
            lds_barrier();
            // waves before this one: lane l < 16 holds wave l's total
            const uint32_t waves = wave_incl_scan(lane < 16 ? S.wsum[lane] : 0u);
            uint32_t run = inc - total;
            if (w) run += (uint32_t)__builtin_amdgcn_readlane((int)waves, (int)w - 1);
            uint32_t e[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                e[k] = run;
                run += c[k];
            }
            uint4* cursor4 = reinterpret_cast<uint4*>(S.u.msd.cursor);
            start4[2 * tid] = cursor4[2 * tid] = make_uint4(e[0], e[1], e[2], e[3]);
            start4[2 * tid + 1] = cursor4[2 * tid + 1] = make_uint4(e[4], e[5], e[6], e[7]);
        }
        const uint32_t largest = wave_max_u32(lane < 16 ? S.wmax[lane] : 0u);
        lds_barrier();
        if (largest <= kBucketCap) {
#pragma unroll
            for (int i = 0; i < kSmallItems; ++i) {
                const uint32_t p = w * 64 * kSmallItems + i * 64 + lane;
                if (p < n) {
                    const uint32_t slot = atomicAdd(&S.u.msd.cursor[key[i] >> kBucketShift], 1u);
                    S.u.msd.comp[slot] = ((key[i] & ((1u << kBucketShift) - 1)) << 12) | p;
                }
            }
            lds_barrier();
            // rank = bucket start + members below; the 4 items' scans interleaved
            uint32_t me[kSmallItems], first[kSmallItems], len[kSmallItems], most = 0;
#pragma unroll
            for (int i = 0; i < kSmallItems; ++i) {
                const uint32_t p = w * 64 * kSmallItems + i * 64 + lane;
                len[i] = 0;
                first[i] = 0;
                me[i] = 0;
                if (p < n) {
                    const uint32_t bkt = key[i] >> kBucketShift;
                    first[i] = S.u.msd.start[bkt];
                    len[i] = S.u.msd.cursor[bkt] - first[i];
                    me[i] = ((key[i] & ((1u << kBucketShift) - 1)) << 12) | p;
                    most = max(most, len[i]);
                }
                rank[i] = first[i];
            }
            // the first kRankUnroll members of every bucket in one batch of
            // independent LDS reads (reads past a bucket are masked; indices
            // stay inside comp), larger buckets continue one member a trip
            constexpr uint32_t kRankUnroll = 8;
#pragma unroll
            for (uint32_t j = 0; j < kRankUnroll; ++j) {
#pragma unroll
                for (int i = 0; i < kSmallItems; ++i) {
                    const uint32_t c = S.u.msd.comp[min(first[i] + j, kSmallMax - 1u)];
                    rank[i] += (j < len[i] && c < me[i]) ? 1u : 0u;
                }
            }
            for (uint32_t j = kRankUnroll; j < most; ++j) {
#pragma unroll
                for (int i = 0; i < kSmallItems; ++i)
                    if (j < len[i]) rank[i] += S.u.msd.comp[first[i] + j] < me[i] ? 1u : 0u;
            }
        } else {
            // LSD: 4 stable passes of 8 bits; padding keys (p >= n) sort last
            uint32_t k2[kSmallItems], val[kSmallItems], dig[kSmallItems], rk[kSmallItems];
            int cur = 0;
#pragma unroll
            for (int i = 0; i < kSmallItems; ++i) {
                k2[i] = key[i];
                val[i] = w * 64 * kSmallItems + i * 64 + lane;
            }
            for (uint32_t shift = 0; shift < 32; shift += 8) {
                if (shift > 0) {
#pragma unroll
                    for (int i = 0; i < kSmallItems; ++i) {
                        const uint32_t p = w * 64 * kSmallItems + i * 64 + lane;
                        k2[i] = S.u.lsd.key[cur][p];
                        val[i] = S.u.lsd.val[cur][p];
                    }
                    cur ^= 1;
                }
#pragma unroll
                for (int i = 0; i < kSmallItems; ++i) dig[i] = (k2[i] >> shift) & 0xFFu;
                block_rank<kSmallBlock, kSmallItems>(dig, rk, S.u.lsd.cnt, S.dstart, S.wsum);
                const int dst = shift == 0 ? 0 : cur;
#pragma unroll
                for (int i = 0; i < kSmallItems; ++i) {
                    S.u.lsd.key[dst][rk[i]] = k2[i];
                    S.u.lsd.val[dst][rk[i]] = val[i];
                }
                cur = dst;
                lds_barrier();
            }
            // rank of triangle p: invert the sorted index array into key[cur]
            for (uint32_t q = tid; q < n; q += kSmallBlock)
                S.u.lsd.key[cur][S.u.lsd.val[cur][q]] = q;
            lds_barrier();
#pragma unroll
            for (int i = 0; i < kSmallItems; ++i) {
                const uint32_t p = w * 64 * kSmallItems + i * 64 + lane;
                rank[i] = p < n ? S.u.lsd.key[cur][p] : 0u;
            }
        }
    }
    lds_barrier();  // sort arrays dead from here
    uint32_t id[kSmallItems];
    bool ok[kSmallItems];
#pragma unroll
    for (int i = 0; i < kSmallItems; ++i) {
        id[i] = w * 64 * kSmallItems + i * 64 + lane;
        ok[i] = id[i] < n;
    }
    small_emit<RADV>(S, n, v, rank, id, key, ok, nodes, tris, ids_out, codes_out, part, parts);
}

__global__ __launch_bounds__(kSmallBlock) void k_build_small(
    const float* __restrict__ pos, const uint32_t* __restrict__ idx, uint32_t n,
    rt_node4* __restrict__ nodes, rt_tri64* __restrict__ tris, uint32_t* __restrict__ ids_out,
    uint32_t* __restrict__ codes_out, SmallTree targ, rt_build_info* info) {
    __shared__ SmallLds S;
    build_small<false>(S, targ, pos, idx, n, nodes, tris, ids_out, codes_out, info, blockIdx.x,
                       gridDim.x);
}

// The RADV image of fast_bvh_builder.comp (header, box32 nodes, 64-B
// triangle nodes) through the fused path; `sort` = 0 keeps the input order
// (the reference builder's leaves of consecutive triangles), 1 the Morton
// order of rt_gpu_build.  tb = total boxes.
__global__ __launch_bounds__(kSmallBlock) void k_build_radv(
    const float* __restrict__ pos, const uint32_t* __restrict__ idx, uint32_t n,
    char* __restrict__ image, uint32_t tb, uint32_t* __restrict__ ids_out, SmallTree targ,
    uint32_t sort) {
    __shared__ SmallLds S;
    build_small<true>(S, targ, pos, idx, n, reinterpret_cast<rt_node4*>(image + 192),
                      reinterpret_cast<rt_tri64*>(image + 192 + 128 * (size_t)tb), ids_out,
                      nullptr, nullptr, blockIdx.x, gridDim.x, sort != 0);
}

// ---------------------------------------------------------------------------
// Medium path: 4096 < n <= 65536 built by ONE workgroup (inside the batched
// build's persistent launch), with the item's scratch in global memory (the
// buffers of the multi-launch path: key / value pairs, level boxes).  The
// steps and the device functions are the multi-launch implicit build's, so
// the output is byte-identical to rt_gpu_build(RT_BUILD_IMPLICIT):
//   1. centroid bounds (wave min / max, then over the 16 waves)
//   2. Morton code and index of every triangle
//   3. four stable LSD passes of 8-bit digits: a digit histogram, its scan,
//      then 4096-key tiles in order, each ranked with block_rank and written
//      to the digit's running position
//   4. sorted ids / codes, leaf records and leaf nodes (emit_leaf_box), then
//      each inner level from the level below (emit_inner), boxes staged in
//      the scratch between levels.
// Global data written by one step and read by another crosses a
// __syncthreads() (workgroup-scope release / acquire: every wave of the
// workgroup runs on this CU).
// ---------------------------------------------------------------------------
constexpr uint32_t kMedMax = 65536;
constexpr int kMedLevels = 8;  // levels of a 65536-triangle implicit tree
constexpr int kMedItems = 4;   // keys per thread and tile
static_assert(RT_BATCH_BUILD_MAX == kMedMax, "the batched build covers the medium path");

struct MedTree {
    uint32_t levels, total_boxes;
    uint32_t counts[kMedLevels];
    uint32_t offsets[kMedLevels];
    uint64_t keys0, vals0, keys1, vals1, boxes;  // byte offsets in the item's scratch
};

// LDS of the medium path: digit counts, running digit starts, the centroid
// bounds, block_rank's per-wave counters, digit starts and wave sums, the
// per-wave partial bounds.
struct MedLdsRefs {
    uint32_t* hist;    // 256
    uint32_t* run;     // 256
    float* bnd;        // 6
    uint32_t* cnt;     // (BLOCK / 64) * 256
    uint32_t* dstart;  // 256
    uint32_t* wsum;    // 16
    float (*red)[6];   // BLOCK / 64 rows
};

// The medium path as its own kernel's LDS (BLOCK = 256: a workgroup that
// shares a compute unit with other kernels' waves).
template <int BLOCK>
struct MedLds {
    uint32_t hist[256];
    uint32_t run[256];
    float bnd[8];
    uint32_t cnt[(BLOCK / 64) * 256];
    uint32_t dstart[256];
    uint32_t wsum[16];
    float red[BLOCK / 64][6];
    __device__ MedLdsRefs refs() { return {hist, run, bnd, cnt, dstart, wsum, red}; }
};

template <int BLOCK>
__device__ void build_medium(const MedLdsRefs& S, const MedTree& T, const float* __restrict__ pos,
                             const uint32_t* __restrict__ idx, uint32_t n,
                             rt_node4* __restrict__ nodes, rt_tri64* __restrict__ tris,
                             uint32_t* __restrict__ ids_out, uint32_t* __restrict__ codes_out,
                             rt_build_info* __restrict__ info, char* __restrict__ scratch) {
    constexpr uint32_t kTile = BLOCK * kMedItems;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    uint32_t* hist = S.hist;  // 256 digit counts
    uint32_t* run = S.run;    // running global start of each digit
    float* bnd = S.bnd;       // lo[3], hi[3]

    // 1. centroid bounds
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t t = tid; t < n; t += BLOCK) {
        float v[9], c[3];
        tri_verts(pos, idx, t, v);
        centroid(v, c);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            mn[a] = fminf(mn[a], c[a]);
            mx[a] = fmaxf(mx[a], c[a]);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float a0 = wave_min(mn[a]), a1 = wave_max(mx[a]);
        if (lane == 0) {
            S.red[w][a] = a0;
            S.red[w][3 + a] = a1;
        }
    }
    lds_barrier();
    if (tid < 6) {
        uint32_t o = f2o(S.red[0][tid]);
        for (int ww = 1; ww < BLOCK / 64; ++ww) {
            const uint32_t q = f2o(S.red[ww][tid]);
            o = tid < 3 ? min(o, q) : max(o, q);
        }
        bnd[tid] = o2f(o);
    }
    lds_barrier();
    float lo[3] = {bnd[0], bnd[1], bnd[2]}, hi[3] = {bnd[3], bnd[4], bnd[5]}, sc[3];
    axis_scale(lo, hi, sc);

    // 2. codes
    uint32_t* ki = reinterpret_cast<uint32_t*>(scratch + T.keys0);
    uint32_t* vi = reinterpret_cast<uint32_t*>(scratch + T.vals0);
    uint32_t* ko = reinterpret_cast<uint32_t*>(scratch + T.keys1);
    uint32_t* vo = reinterpret_cast<uint32_t*>(scratch + T.vals1);
    for (uint32_t t = tid; t < n; t += BLOCK) {
        float v[9], c[3];
        tri_verts(pos, idx, t, v);
        centroid(v, c);
        ki[t] = morton(c, lo, sc);
        vi[t] = t;
    }
    __syncthreads();

    // 3. LSD passes (four: the codes end in the first buffers)
    for (uint32_t shift = 0; shift < 32; shift += 8) {
        if (tid < 256) hist[tid] = 0;
        lds_barrier();
        for (uint32_t t = tid; t < n; t += BLOCK)
            atomicAdd(&hist[(ki[t] >> shift) & 0xFFu], 1u);
        lds_barrier();
        uint32_t hv[1] = {tid < 256 ? hist[tid] : 0u};
        digit_excl_scan<256>(hv, S.wsum);
        if (tid < 256) run[tid] = hv[0];
        lds_barrier();
        for (uint32_t base = 0; base < n; base += kTile) {
            const uint32_t valid = min(kTile, n - base);
            uint32_t key[kMedItems], val[kMedItems], dig[kMedItems], rk[kMedItems];
#pragma unroll
            for (int i = 0; i < kMedItems; ++i) {
                const uint32_t lp = w * 64 * kMedItems + i * 64 + lane;
                key[i] = lp < valid ? ki[base + lp] : 0xFFFFFFFFu;
                val[i] = lp < valid ? vi[base + lp] : 0u;
                dig[i] = (key[i] >> shift) & 0xFFu;
            }
            block_rank<BLOCK, kMedItems>(dig, rk, S.cnt, S.dstart, S.wsum);
#pragma unroll
            for (int i = 0; i < kMedItems; ++i) {
                const uint32_t lp = w * 64 * kMedItems + i * 64 + lane;
                if (lp < valid) {
                    const uint32_t g = run[dig[i]] + rk[i] - S.dstart[dig[i]];
                    ko[g] = key[i];
                    vo[g] = val[i];
                }
            }
            lds_barrier();
            // advance each digit by its count in this tile (the padding keys
            // all have digit 255 and rank after the valid ones)
            if (tid < 256) {
                const uint32_t next = tid < 255 ? S.dstart[tid + 1] : valid;
                run[tid] += next - S.dstart[tid];
            }
            lds_barrier();
        }
        __syncthreads();
        uint32_t* t0 = ki; ki = ko; ko = t0;
        t0 = vi; vi = vo; vo = t0;
    }

    // 4. outputs
    for (uint32_t q = tid; q < n; q += BLOCK) {
        if (ids_out) ids_out[q] = vi[q];
        if (codes_out) codes_out[q] = ki[q];
    }
    float* gb = reinterpret_cast<float*>(scratch + T.boxes);
    const uint32_t lb = T.levels - 1;
    for (uint32_t b = tid; b < T.counts[lb]; b += BLOCK) {
        float box[6];
        emit_leaf_box<false>(pos, idx, n, vi, nodes, tris, T.offsets[lb], b, box);
#pragma unroll
        for (int a = 0; a < 6; ++a) gb[6 * (size_t)(T.offsets[lb] + b) + a] = box[a];
    }
    __syncthreads();
    for (int lv = (int)lb - 1; lv >= 0; --lv) {
        for (uint32_t i = tid; i < T.counts[lv]; i += BLOCK) {
            float box[6];
            emit_inner(nodes, T, (uint32_t)lv, i, gb + 6 * (size_t)T.offsets[lv + 1], 0u, box);
            if (lv > 0) {
#pragma unroll
                for (int a = 0; a < 6; ++a) gb[6 * (size_t)(T.offsets[lv] + i) + a] = box[a];
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        info->node_count = T.total_boxes;
        info->depth = T.levels;
    }
}

// ---------------------------------------------------------------------------
// Generic refit of any Box4 tree whose inner children have larger indices
// than their parents (radix pre-order, implicit level order, an uploaded
// SoftwareBVH4): topology kept, records and boxes from new positions.
//   k_refit_leaf_slots  thread per node: every leaf-range slot rewrites its
//                       records (vertices re-read by triangle_id) and gets
//                       the union of their boxes; the node's round is reset.
//   k_refit_round(r)    thread per unfinished node: if every inner child
//                       finished in an earlier round (done < r), the node's
//                       inner slots get the union of the child's occupied
//                       slots and done = r.  A node of height h finishes in
//                       round h, so `depth` rounds finish the root.  Reads of
//                       a flag written in the same round see either the old
//                       or the new value, and neither is < r: the kernel
//                       boundaries order every box a node reads.
// ---------------------------------------------------------------------------
constexpr uint32_t kNotDone = 0xFFFFFFFFu;

__device__ __forceinline__ void set_slot_box(rt_node4* nd, int k, const float b[6]) {
    nd->minX[k] = b[0];
    nd->minY[k] = b[1];
    nd->minZ[k] = b[2];
    nd->maxX[k] = b[3];
    nd->maxY[k] = b[4];
    nd->maxZ[k] = b[5];
}

__global__ __launch_bounds__(256) void k_refit_leaf_slots(const float* __restrict__ pos,
                                                          const uint32_t* __restrict__ idx,
                                                          rt_node4* __restrict__ nodes,
                                                          uint32_t node_count,
                                                          rt_tri64* __restrict__ tris,
                                                          uint32_t* __restrict__ done) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= node_count) return;
    done[i] = kNotDone;
    const int4 ch = reinterpret_cast<const int4*>(nodes + i)[6];
    const int4 cn = reinterpret_cast<const int4*>(nodes + i)[7];
    const int32_t chv[4] = {ch.x, ch.y, ch.z, ch.w};
    const int32_t cnv[4] = {cn.x, cn.y, cn.z, cn.w};
    for (int k = 0; k < 4; ++k) {
        if (cnv[k] <= 0) continue;
        float u[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        const uint32_t first = (uint32_t)~chv[k];
        for (int32_t q = 0; q < cnv[k]; ++q) {
            const uint32_t s = first + (uint32_t)q;
            const uint32_t id = tris[s].triangle_id;
            float v[9], bb[6];
            tri_verts(pos, idx, id, v);
            store_tri(tris + s, v, id);
            tri_box(v, bb);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                u[a] = fminf(u[a], bb[a]);
                u[3 + a] = fmaxf(u[3 + a], bb[3 + a]);
            }
        }
        set_slot_box(nodes + i, k, u);
    }
}

__global__ __launch_bounds__(256) void k_refit_round(rt_node4* __restrict__ nodes,
                                                     uint32_t node_count,
                                                     uint32_t* __restrict__ done, uint32_t r) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= node_count || done[i] != kNotDone) return;
    const int4 ch = reinterpret_cast<const int4*>(nodes + i)[6];
    const int4 cn = reinterpret_cast<const int4*>(nodes + i)[7];
    const int32_t chv[4] = {ch.x, ch.y, ch.z, ch.w};
    const int32_t cnv[4] = {cn.x, cn.y, cn.z, cn.w};
    for (int k = 0; k < 4; ++k) {
        if (cnv[k] >= 0) continue;
        const uint32_t d = done[chv[k]];
        if (d == kNotDone || d >= r) return;  // a child not finished before this round
    }
    for (int k = 0; k < 4; ++k) {
        if (cnv[k] >= 0) continue;
        const rt_node4& c = nodes[chv[k]];
        float u[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        for (int q = 0; q < 4; ++q) {
            if (c.count[q] == 0) continue;
            u[0] = fminf(u[0], c.minX[q]);
            u[1] = fminf(u[1], c.minY[q]);
            u[2] = fminf(u[2], c.minZ[q]);
            u[3] = fmaxf(u[3], c.maxX[q]);
            u[4] = fmaxf(u[4], c.maxY[q]);
            u[5] = fmaxf(u[5], c.maxZ[q]);
        }
        set_slot_box(nodes + i, k, u);
    }
    done[i] = r;
}

// Device table entry of the batched build: the caller's item and the level
// table of its tree (fused path for n <= 4096, medium path above).
struct BatchBuild {
    rt_build_item item;
    SmallTree tree;
    MedTree med;
};

// Persistent batched build: workgroup b builds items b, b + grid, b + 2 grid,
// ... (the host orders them largest first), each alone (parts = 1).
// skip_medium: items above kSmallMax are left to k_build_batch_shared.
__global__ __launch_bounds__(kSmallBlock) void k_build_batch(const BatchBuild* __restrict__ table,
                                                              uint32_t count, uint32_t skip_medium) {
    __shared__ SmallLds S;
    const MedLdsRefs M{S.u.lsd.key[0], S.u.lsd.key[0] + 256,
                       reinterpret_cast<float*>(S.u.lsd.val[0]), S.u.lsd.cnt, S.dstart, S.wsum,
                       S.red};
    for (uint32_t k = blockIdx.x; k < count; k += gridDim.x) {
        const BatchBuild& b = table[k];
        const rt_build_item& it = b.item;
        if (it.n <= kSmallMax)
            build_small<false>(S, b.tree, it.pos, it.idx, it.n, it.nodes, it.tris, it.sorted_ids,
                               it.codes, it.info, 0u, 1u);
        else if (!skip_medium)
            build_medium<kSmallBlock>(M, b.med, it.pos, it.idx, it.n, it.nodes, it.tris,
                                      it.sorted_ids, it.codes, it.info,
                                      static_cast<char*>(it.scratch));
        __syncthreads();  // LDS and scratch reused by the next mesh
    }
}

// The medium items of a batch (RT_BUILD_BATCH_SHARED): the same medium path
// by 256-thread workgroups with about 7 KB of LDS, which fit on a compute
// unit beside other kernels' waves (a 1024-thread workgroup with the fused
// path's LDS needs an idle compute unit, and waits for one while traces
// run).  Workgroup b builds items b, b + grid, ...; items of kSmallMax
// triangles or fewer are skipped (k_build_batch builds them).
constexpr int kSharedBlock = 256;
__global__ __launch_bounds__(kSharedBlock) void k_build_batch_shared(
    const BatchBuild* __restrict__ table, uint32_t count) {
    __shared__ MedLds<kSharedBlock> S;
    const MedLdsRefs M = S.refs();
    for (uint32_t k = blockIdx.x; k < count; k += gridDim.x) {
        const BatchBuild& b = table[k];
        const rt_build_item& it = b.item;
        if (it.n <= kSmallMax) continue;
        build_medium<kSharedBlock>(M, b.med, it.pos, it.idx, it.n, it.nodes, it.tris,
                                   it.sorted_ids, it.codes, it.info,
                                   static_cast<char*>(it.scratch));
        __syncthreads();  // LDS and scratch reused by the next mesh
    }
}

// Refit of a fused-path tree: sorted position s = i*1024 + t keeps its
// triangle (the index in its record); vertices are re-read and every box
// and record rewritten through the same output path.  Every workgroup reads
// every record's triangle index while others rewrite their slices; the
// index is rewritten with its own value, so a read sees it either way.
__global__ __launch_bounds__(kSmallBlock) void k_refit_small(
    const float* __restrict__ pos, const uint32_t* __restrict__ idx, uint32_t n,
    rt_node4* __restrict__ nodes, rt_tri64* __restrict__ tris, SmallTree targ,
    rt_build_info* info) {
    __shared__ SmallLds S;
    small_info(info, targ, blockIdx.x);
    load_levels(S.lv, targ);  // first read after small_emit's first barrier
    float v[kSmallItems][9];
    uint32_t r[kSmallItems], id[kSmallItems], key[kSmallItems];
    bool ok[kSmallItems];
#pragma unroll
    for (int i = 0; i < kSmallItems; ++i) {
        r[i] = i * kSmallBlock + threadIdx.x;
        ok[i] = r[i] < n;
        id[i] = ok[i] ? tris[r[i]].triangle_id : 0u;
        key[i] = 0;
    }
#pragma unroll
    for (int i = 0; i < kSmallItems; ++i)
        if (ok[i]) tri_verts(pos, idx, id[i], v[i]);
    small_emit<false>(S, n, v, r, id, key, ok, nodes, tris, nullptr, nullptr, blockIdx.x,
                      gridDim.x);
}

// ---------------------------------------------------------------------------
// Radix-tree path: binary radix tree over the sorted codes (Karras 2012,
// equal codes ordered by sorted position), ranges of <= maxp triangles are
// leaves, collapsed to Box4 with the greedy largest-surface-area expansion of
// software_bvh.hpp:225-247 and numbered in the pre-order of
// software_bvh.hpp:249-277.
//
// Node ids: internal node i in [0, m), leaf (sorted position) p -> m + p,
// m = n - 1.  Internal node i covers sorted range [rf[i], rl[i]].
// ---------------------------------------------------------------------------
// Sorted positions per bottom-up window (one workgroup, LDS counters).
constexpr uint32_t kWindow = 4096;

struct RadixArrays {
    const uint32_t* keys;  // sorted codes
    float* box;            // (2n - 1) x 6: min xyz, max xyz
    uint32_t* rf;          // m
    uint32_t* rl;          // m
    int32_t* left;         // m
    int32_t* right;        // m
    int32_t* parent;       // 2n - 1
    uint32_t* done;        // m: arrivals at a straddling node (k_rt_top)
    uint32_t* top;         // straddling nodes (range crosses a kWindow boundary)
    uint32_t* top_count;   // 1 word
    uint32_t* starts;      // n / kWindow + 1: k_rt_top climb starts
    int4* cand;            // m, -1 = empty
    uint32_t* size4;       // m: Box4 nodes in the subtree rooted here
    float* cost;           // m: SAH cost of the subtree (treelet restructuring)
    uint32_t* emin;        // m: lowest leaf rank below the node (see k_tl_rank_jump)
    uint32_t* emax;        // m: highest leaf rank below the node
    // treelet restructuring only (null otherwise)
    uint32_t* cnt;         // m: triangles below the node
    uint64_t* jump;        // m: (ancestor, rank offset) pairs of the rank pass
    uint32_t* rank;        // n: leaf rank of sorted position p (tree order)
    uint32_t* order;       // n: sorted position of leaf rank r
    uint32_t n, m, maxp;
    uint32_t wshift;       // treelet passes: windows of 1 << wshift leaf ranks
};

// Highest differing bit of sorted positions i and j: key bits, then 32 +
// position bits; -1 outside [0, n).
__device__ __forceinline__ int rdelta(const uint32_t* __restrict__ k, uint32_t n, int64_t i,
                                      int64_t j) {
    if (j < 0 || j >= (int64_t)n) return -1;
    const uint32_t x = k[i] ^ k[j];
    if (x) return __clz(x);
    return 32 + __clz((uint32_t)(i ^ j));
}

__device__ __forceinline__ uint32_t nfirst(const RadixArrays& A, int32_t u) {
    return (uint32_t)u < A.m ? A.rf[u] : (uint32_t)u - A.m;
}
__device__ __forceinline__ uint32_t nlast(const RadixArrays& A, int32_t u) {
    return (uint32_t)u < A.m ? A.rl[u] : (uint32_t)u - A.m;
}
__device__ __forceinline__ bool ninner(const RadixArrays& A, int32_t u) {
    return (uint32_t)u < A.m && A.rl[u] - A.rf[u] + 1 > A.maxp;
}

// Sorted position p: 64-B leaf record and its box (canonical zeros).
__global__ __launch_bounds__(256) void k_rt_leaves(const float* __restrict__ pos,
                                                   const uint32_t* __restrict__ idx, uint32_t n,
                                                   const uint32_t* __restrict__ sorted,
                                                   rt_tri64* __restrict__ tris,
                                                   float* __restrict__ box, uint32_t m,
                                                   uint32_t* __restrict__ top_count) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p == 0) *top_count = 0;  // appended to by k_rt_radix
    if (p >= n) return;
    const uint32_t id = sorted[p];
    float v[9];
    tri_verts(pos, idx, id, v);
    store_tri(tris + p, v, id);
    float bb[6];
    tri_box(v, bb);
    float* b = box + 6 * (size_t)(m + p);
#pragma unroll
    for (int a = 0; a < 6; ++a) b[a] = bb[a];
}

// Internal node i (Karras 2012, Alg. 1 with the position tie-break).
__global__ __launch_bounds__(256) void k_rt_radix(RadixArrays A, rt_build_info* info) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= A.m) return;
    const uint32_t* k = A.keys;
    const uint32_t n = A.n;
    const int64_t ii = i;
    const int d = (rdelta(k, n, ii, ii + 1) - rdelta(k, n, ii, ii - 1)) >= 0 ? 1 : -1;
    const int dmin = rdelta(k, n, ii, ii - d);
    int64_t lmax = 2;
    while (rdelta(k, n, ii, ii + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (rdelta(k, n, ii, ii + (l + t) * d) > dmin) l += t;
    const int64_t j = ii + l * d;
    const int dnode = rdelta(k, n, ii, j);
    int64_t s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (rdelta(k, n, ii, ii + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int64_t g = ii + s * d + (d < 0 ? -1 : 0);
    const int64_t lo = ii < j ? ii : j, hi = ii < j ? j : ii;
    const int32_t L = (lo == g) ? (int32_t)(A.m + g) : (int32_t)g;
    const int32_t R = (hi == g + 1) ? (int32_t)(A.m + g + 1) : (int32_t)(g + 1);
    A.rf[i] = (uint32_t)lo;
    A.rl[i] = (uint32_t)hi;
    A.emin[i] = (uint32_t)lo;
    A.emax[i] = (uint32_t)hi;
    A.left[i] = L;
    A.right[i] = R;
    A.parent[L] = (int32_t)i;
    A.parent[R] = (int32_t)i;
    A.done[i] = 0;
    if (lo / kWindow != hi / kWindow)  // finished by k_rt_top, not by a window
        A.top[atomicAdd(A.top_count, 1u)] = i;
    if (i == 0) {
        A.parent[0] = -1;
        info->node_count = 0;
        info->depth = 0;
    }
}

__device__ __forceinline__ float box_sa(const float* b) {
    // software_bvh.hpp:231-234
    const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// The greedy Box4 candidates and the Box4 subtree size of internal node x,
// once the boxes of its subtree and the sizes below it are final.
__device__ void finish_box4(const RadixArrays& A, int32_t x) {
    if (A.rl[x] - A.rf[x] + 1 > A.maxp) {
        int32_t c[4] = {A.left[x], A.right[x], -1, -1};
        int nc = 2;
        while (nc < 4) {  // software_bvh.hpp:225-247
            int best = -1;
            float bestSa = -1.0f;
            for (int q = 0; q < nc; ++q) {
                if (!ninner(A, c[q])) continue;
                const float sa = box_sa(A.box + 6 * (size_t)c[q]);
                if (sa > bestSa) {
                    bestSa = sa;
                    best = q;
                }
            }
            if (best < 0) break;
            const int32_t e = c[best];
            c[best] = A.left[e];
            c[nc++] = A.right[e];
        }
        uint32_t sz = 1;
        for (int q = 0; q < nc; ++q)
            if (ninner(A, c[q])) sz += A.size4[c[q]];
        A.cand[x] = make_int4(c[0], c[1], c[2], c[3]);
        A.size4[x] = sz;
    }
}

// Finish internal node x once both subtrees are: box = union of the
// children, then the Box4 candidates and size.
__device__ void finish_node(const RadixArrays& A, int32_t x) {
    const int32_t L = A.left[x], R = A.right[x];
    float* bx = A.box + 6 * (size_t)x;
    const float* bl = A.box + 6 * (size_t)L;
    const float* br = A.box + 6 * (size_t)R;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        bx[a] = fminf(bl[a], br[a]);
        bx[3 + a] = fmaxf(bl[3 + a], br[3 + a]);
    }
    finish_box4(A, x);
}

__device__ __forceinline__ bool straddles(const RadixArrays& A, int32_t x) {
    return (uint32_t)x < A.m && A.rf[x] / kWindow != A.rl[x] / kWindow;
}

// Pass 1: block b owns the kWindow sorted positions [b kWindow, ...); each
// leaf climbs and, at every internal node whose range lies inside the
// window, the second thread to arrive (LDS counter, workgroup-scope
// acquire-release) finishes it.  A climb stops at the first node that
// straddles a window boundary: those are finished by k_rt_top.  (Agent-scope
// release writes back this XCD's L2; one such atomic per node made this pass
// ~10 ms at 2.5 M triangles, so no agent-scope atomic is used at all.)
__global__ __launch_bounds__(1024) void k_rt_bottom_up(RadixArrays A) {
    __shared__ uint32_t wcnt[kWindow];
    const uint32_t lo = blockIdx.x * kWindow;
    const uint32_t hi = min(lo + kWindow, A.n) - 1;
    for (uint32_t i = threadIdx.x; i < kWindow; i += 1024) wcnt[i] = 0;
    __syncthreads();
    for (uint32_t p = lo + threadIdx.x; p <= hi; p += 1024) {
        int32_t x = A.parent[A.m + p];
        while (x >= 0) {
            // x covers [rf[x], rl[x]] with x at one end, so an inside node's
            // counter slot x - lo is in [0, kWindow)
            if (!(A.rf[x] >= lo && A.rl[x] <= hi)) break;
            const uint32_t old = __hip_atomic_fetch_add(wcnt + (x - lo), 1u, __ATOMIC_ACQ_REL,
                                                        __HIP_MEMORY_SCOPE_WORKGROUP);
            if (old == 0) break;  // the sibling subtree is still in flight
            finish_node(A, x);
            x = A.parent[x];
        }
    }
}

// Pass 2, one workgroup: the straddling nodes (at most depth per window
// boundary; their parents straddle too).  A child that does not straddle was
// finished by pass 1 and counts as an arrival up front; a straddling child
// arrives when its own climb finishes it; the second arrival finishes the
// node and climbs on.  Climbs start at the nodes whose children both do not
// straddle (one per boundary at most), gathered first so that no thread
// holds two of them.  The arrival counters are in global memory, but every
// access is from this one workgroup, so workgroup scope suffices.
__global__ __launch_bounds__(1024) void k_rt_top(RadixArrays A) {
    __shared__ uint32_t nstart;
    const uint32_t cnt = *A.top_count;
    if (threadIdx.x == 0) nstart = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < cnt; i += 1024) {
        const int32_t x = (int32_t)A.top[i];
        const int ready = (straddles(A, A.left[x]) ? 0 : 1) + (straddles(A, A.right[x]) ? 0 : 1);
        if (ready == 1)
            A.done[x] = 1;  // no climb has started yet: plain store
        else if (ready == 2)
            A.starts[atomicAdd(&nstart, 1u)] = (uint32_t)x;
    }
    __syncthreads();
    const uint32_t ns = nstart;
    for (uint32_t j = threadIdx.x; j < ns; j += 1024) {
        int32_t x = (int32_t)A.starts[j];
        for (;;) {
            finish_node(A, x);
            x = A.parent[x];
            if (x < 0) break;
            if (__hip_atomic_fetch_add(A.done + x, 1u, __ATOMIC_ACQ_REL,
                                       __HIP_MEMORY_SCOPE_WORKGROUP) == 0)
                break;
        }
    }
}

// ---------------------------------------------------------------------------
// Treelet restructuring of the radix tree (Karras & Aila 2013, "Fast Parallel
// Construction of High-Quality Bounding Volume Hierarchies"), bottom-up.  At
// each inner node x (more than maxp triangles) a treelet of up to 7 leaves
// grows from x's two children by expanding the inner treelet leaf of largest
// surface area (the first in slot order on ties).  A dynamic program over the
// leaf subsets finds the binary tree over those leaves of least SAH cost
//   leaf range: SA * triangles;  inner node: 1.2 * SA + cost(left) + cost(right)
// (each split counted once: the part holding the subset's lowest leaf, scanned
// from s - 1 down, the first strict minimum kept).  When that cost is strictly
// below the current one the treelet is rebuilt, reusing its inner nodes in
// breadth-first order.  Leaf ranges, and so the triangle order, never change;
// inner nodes keep their radix ranges in rf / rl (which then only say that
// they are inner), and emin / emax follow the sorted positions below them.
// A node's result depends only on its subtree, so the order in which threads
// reach the nodes does not change the tree.
// ---------------------------------------------------------------------------
constexpr int kTreeletLeaves = 7;
constexpr uint32_t kMaxTreeletRounds = 16;
constexpr uint32_t kTreeletSets = 1u << kTreeletLeaves;
// 4 waves; 3 blocks fit a CU (registers, and LDS: 52,864 bytes a block; at
// 53,888 only 2 fit, which made a round over 2.5 M triangles 1.6x slower)
constexpr uint32_t kTreeletBlock = 256;
constexpr uint32_t kTreeletTopBlocks = 1024;  // pass 2 grid (walks the start list)
constexpr float kCostInner = 1.2f;

__device__ __forceinline__ uint32_t ext_lo(const RadixArrays& A, int32_t u) {
    return (uint32_t)u < A.m ? A.emin[u] : A.rank[(uint32_t)u - A.m];
}
__device__ __forceinline__ uint32_t ext_hi(const RadixArrays& A, int32_t u) {
    return (uint32_t)u < A.m ? A.emax[u] : A.rank[(uint32_t)u - A.m];
}
// Triangles below u.
__device__ __forceinline__ uint32_t ncount(const RadixArrays& A, int32_t u) {
    return (uint32_t)u < A.m ? A.cnt[u] : 1u;
}
__device__ __forceinline__ bool ext_straddles(const RadixArrays& A, int32_t x) {
    return (uint32_t)x < A.m && A.emin[x] >> A.wshift != A.emax[x] >> A.wshift;
}

// SAH cost of the subtree at u: a leaf range SA * triangles, an inner node its
// stored cost.
__device__ __forceinline__ float sah_cost(const RadixArrays& A, int32_t u) {
    if (ninner(A, u)) return A.cost[u];
    return box_sa(A.box + 6 * (size_t)u) * (float)(nlast(A, u) - nfirst(A, u) + 1);
}

// Cheapest split of leaf subset s over the cost table C (C[t * STRIDE] =
// cost of subset t).  The parts holding s's lowest leaf, in descending order,
// are low | sub for the proper submasks sub of rest = s ^ low, descending; four
// are loaded at a time and compared in order, the first strict minimum kept.
template <uint32_t STRIDE>
__device__ __forceinline__ float best_split(const float* C, uint32_t s, uint32_t& bp) {
    const uint32_t low = s & (0u - s), rest = s ^ low;
    float best = INFINITY;
    bp = 0;
    uint32_t sub = (rest - 1) & rest;
    bool end = false;
    while (!end) {
        uint32_t q[4];
        int k = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (end) continue;
            q[j] = sub;
            k = j + 1;
            if (sub == 0)
                end = true;
            else
                sub = (sub - 1) & rest;
        }
        float a[4], b[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j >= k) continue;
            a[j] = C[(low | q[j]) * STRIDE];
            b[j] = C[(rest ^ q[j]) * STRIDE];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j >= k) continue;
            const float c = a[j] + b[j];
            if (c < best) {
                best = c;
                bp = low | q[j];
            }
        }
    }
    return best;
}

// LDS ordering between the steps of a wave's dynamic program: the wave's own
// LDS writes complete before its later reads (LDS instructions of one wave
// execute in order); this keeps the compiler from moving them across.
__device__ __forceinline__ void wave_lds_step() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One wave's LDS: the subset cost and split tables of the treelet in
// progress, and a row per lane with the leaf boxes (7 x 6 floats) and leaf costs (7) of the
// treelet that lane formed.
constexpr uint32_t kRowFloats = 6 * kTreeletLeaves + kTreeletLeaves;  // 49, unpadded
struct TlShared {
    float C[kTreeletSets];
    uint8_t P[kTreeletSets];  // the split best_split chooses for each subset (0: none)
    float row[64 * kRowFloats];
};

// A treelet as formed by one lane: leaves lf[0..nl), inner nodes
// in[0..nl-1) (in[0] = the root x, then in the order they were expanded),
// the root's current cost.  nl = 0: the root is a leaf range (no step);
// nl = 2: neither child is inner (nothing to restructure).
struct TlForm {
    int32_t lf[kTreeletLeaves];
    int32_t in[kTreeletLeaves - 1];
    int32_t nl;
    float cx;
};

// The subsets of 7 leaves by size, ascending within a size (so the size-k
// subsets of nl < 7 leaves are the first C(nl, k) of size k), where each
// size starts, and the binomial coefficients C(nl, k).  The treelet kernels
// copy the subset list to LDS (lanes index it by their group); the other two
// are read with wave-uniform indices (scalar loads: 32-bit entries).
__constant__ uint8_t kSubsetsBySize[kTreeletSets] = {
    0,  1,  2,  4,  8,  16, 32, 64, 3,  5,  6,  9,  10, 12, 17, 18, 20, 24, 33, 34, 36, 40,
    48, 65, 66, 68, 72, 80, 96, 7,  11, 13, 14, 19, 21, 22, 25, 26, 28, 35, 37, 38, 41, 42,
    44, 49, 50, 52, 56, 67, 69, 70, 73, 74, 76, 81, 82, 84, 88, 97, 98, 100, 104, 112, 15, 23,
    27, 29, 30, 39, 43, 45, 46, 51, 53, 54, 57, 58, 60, 71, 75, 77, 78, 83, 85, 86, 89, 90,
    92, 99, 101, 102, 105, 106, 108, 113, 114, 116, 120, 31, 47, 55, 59, 61, 62, 79, 87, 91, 93,
    94, 103, 107, 109, 110, 115, 117, 118, 121, 122, 124, 63, 95, 111, 119, 123, 125, 126, 127};
__constant__ uint32_t kSubsetStart[kTreeletLeaves + 1] = {0, 1, 8, 29, 64, 99, 120, 127};
__constant__ uint32_t kBinom[kTreeletLeaves + 1][kTreeletLeaves + 1] = {
    {1, 0, 0, 0, 0, 0, 0, 0},     {1, 1, 0, 0, 0, 0, 0, 0},      {1, 2, 1, 0, 0, 0, 0, 0},
    {1, 3, 3, 1, 0, 0, 0, 0},     {1, 4, 6, 4, 1, 0, 0, 0},      {1, 5, 10, 10, 5, 1, 0, 0},
    {1, 6, 15, 20, 15, 6, 1, 0},  {1, 7, 21, 35, 35, 21, 7, 1}};

// The bits of v deposited, lowest first, at the set bits of mask: the
// proper submasks of mask in descending order are v = 2^popc(mask) - 2 ... 0.
__device__ __forceinline__ uint32_t deposit_bits(uint32_t v, uint32_t mask) {
    uint32_t r = 0;
    for (uint32_t b = mask; b; b &= b - 1u, v >>= 1)
        if (v & 1u) r |= b & (0u - b);
    return r;
}

// Surface area of u when it is an inner node (a candidate to expand), else -1.
__device__ __forceinline__ float expand_sa(const RadixArrays& A, int32_t u) {
    return ninner(A, u) ? box_sa(A.box + 6 * (size_t)u) : -1.0f;
}

// Treelet formation at x by one lane (Karras & Aila 2013 grow the treelet
// from x's children, expanding the inner leaf of largest surface area, the
// first in slot order on ties), with the leaves' boxes and costs written to
// the lane's LDS row.  Lanes form their own nodes' treelets side by side, so
// the dependent loads of one formation overlap those of the others.
__device__ void treelet_form(const RadixArrays& A, int32_t x, TlForm& f, float* row) {
    f.nl = 0;
    if (!ninner(A, x)) return;
    f.lf[0] = A.left[x];
    f.lf[1] = A.right[x];
    f.in[0] = x;
    float sa[kTreeletLeaves];
    sa[0] = expand_sa(A, f.lf[0]);
    sa[1] = expand_sa(A, f.lf[1]);
    f.cx = kCostInner * box_sa(A.box + 6 * (size_t)x) +
           (sah_cost(A, f.lf[0]) + sah_cost(A, f.lf[1]));
    int nl = 2;
    // unrolled: expansion `it` has it + 2 leaves (static slots, so the
    // arrays stay in registers)
#pragma unroll
    for (int it = 0; it < kTreeletLeaves - 2; ++it) {
        const int have = it + 2;
        int best = -1;
        float bs = -1.0f;
#pragma unroll
        for (int q = 0; q < have; ++q)
            if (sa[q] > bs) {  // -1 (a leaf range) never is
                bs = sa[q];
                best = q;
            }
        if (best < 0) break;
        int32_t e = 0;
#pragma unroll
        for (int q = 0; q < have; ++q)
            if (q == best) e = f.lf[q];
        f.in[it + 1] = e;
        const int32_t l = A.left[e], r = A.right[e];
        const float sl = expand_sa(A, l), sr = expand_sa(A, r);
#pragma unroll
        for (int q = 0; q < have; ++q)
            if (q == best) {
                f.lf[q] = l;
                sa[q] = sl;
            }
        f.lf[have] = r;
        sa[have] = sr;
        nl = have + 1;
    }
    f.nl = nl;
    if (nl < 3) return;
#pragma unroll
    for (int q = 0; q < kTreeletLeaves; ++q) {
        if (q >= nl) continue;
        const float* b = A.box + 6 * (size_t)f.lf[q];
#pragma unroll
        for (int a = 0; a < 6; ++a) row[6 * q + a] = b[a];
        row[6 * kTreeletLeaves + q] = sah_cost(A, f.lf[q]);
    }
}

// The dynamic program and rebuild of one formed treelet (root x, wave-
// uniform arguments: lf, in, nl, cx read from the forming lane; row = that
// lane's LDS row), by the whole wave (Karras & Aila 2013 process one treelet
// per warp): subset costs go to the wave's table C, the subsets of one size
// at a time, lane l holding subsets l + 1 and l + 65 (its own scan of their
// splits, as best_split, its choice kept in P); the full set's splits are
// spread over the lanes, in best_split's order, and reduced to their least
// cost and the first lane reaching it; lane 0 rebuilds the treelet from P
// when that is strictly cheaper and stores x's cost.  The costs, the splits
// chosen and so the tree are those of the sequential program (each subset's
// cost is a least value, which does not depend on the order its candidates
// are compared in; each choice is the first strict minimum in best_split's
// order).
// A rebuilt inner node's box is the union of its subset's leaf boxes (exact:
// min and max do not round), its rank range and count those of the leaves;
// the root's box, range and count do not change (same leaves).
__device__ void treelet_step(const RadixArrays& A, int32_t x, int nl, float cx,
                             const int32_t* lf, const int32_t* in, const float* row,
                             TlShared* sh, const uint8_t* subsets) {
    const uint32_t lane = threadIdx.x & 63u;
    float* C = sh->C;
    uint8_t* P = sh->P;
    if (nl >= 3) {
        // lane q < nl fetches leaf q's rank range and count now, for a
        // rebuild at the end (read there by lane 0)
        uint32_t mlo = 0, mhi = 0, mc = 0;
        if (lane < (uint32_t)nl) {
            int32_t u = lf[0];  // lf[lane], by masks (no indexed private array)
#pragma unroll
            for (int q = 1; q < kTreeletLeaves; ++q)
                u ^= (u ^ lf[q]) & -(int32_t)(lane == (uint32_t)q);
            mlo = ext_lo(A, u);
            mhi = ext_hi(A, u);
            mc = ncount(A, u);
        }
        wave_lds_step();  // the previous treelet's reads of C are done
        float lb[kTreeletLeaves][6];
#pragma unroll
        for (int q = 0; q < kTreeletLeaves; ++q) {
            if (q >= nl) continue;
#pragma unroll
            for (int a = 0; a < 6; ++a) lb[q][a] = row[6 * q + a];
        }
        if (lane < (uint32_t)nl) C[1u << lane] = row[6 * kTreeletLeaves + lane];
        const uint32_t full = (1u << nl) - 1;
        // surface areas of the subsets of two leaves or more (as the
        // sequential union), parked in C until the subset's cost replaces it
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t sl = lane + 1u + 64u * h;
            float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
#pragma unroll
            for (int q = 0; q < kTreeletLeaves; ++q) {
                if (!((sl >> q) & 1u)) continue;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    b[a] = fminf(b[a], lb[q][a]);
                    b[3 + a] = fmaxf(b[3 + a], lb[q][3 + a]);
                }
            }
            if (sl <= full && (sl & (sl - 1))) C[sl] = box_sa(b);
        }
        wave_lds_step();
        // subsets by size k: the S size-k subsets (kSubsetsBySize) each get a
        // group of G lanes (G the largest power of two with G S <= 64), which
        // share its 2^(k-1) - 1 splits in best_split's order, and reduce
        // their first strict minima to the least cost, then the first split
        for (int k = 2; k < nl; ++k) {
            const uint32_t S = kBinom[nl][k];  // 3 <= S <= 35
            const uint32_t lg = __clz(S - 1u) - 26u;  // 6 - ceil(log2 S)
            const uint32_t t = lane >> lg, g = lane & ((1u << lg) - 1u);
            const uint32_t ncand = (1u << (k - 1)) - 1u;
            const uint32_t per = (ncand + (1u << lg) - 1u) >> lg;
            uint32_t s = 0, low = 0, rest = 0;
            float best = INFINITY;
            uint32_t bj = 0xffffffffu;
            if (t < S) {
                s = subsets[kSubsetStart[k] + t];
                low = s & (0u - s);
                rest = s ^ low;
                const uint32_t j1 = min(g * per + per, ncand);
                uint32_t sub = deposit_bits(ncand - 1u - g * per, rest);
                for (uint32_t j = g * per; j < j1; ++j) {
                    const float c = C[low | sub] + C[rest ^ sub];
                    if (c < best) {
                        best = c;
                        bj = j;
                    }
                    sub = (sub - 1u) & rest;
                }
            }
            for (uint32_t o = 1; o < (1u << lg); o <<= 1) {
                const float ob = __shfl_xor(best, (int)o);
                const uint32_t oj = (uint32_t)__shfl_xor((int)bj, (int)o);
                if (ob < best || (ob == best && oj < bj)) {
                    best = ob;
                    bj = oj;
                }
            }
            if (t < S && g == 0) {
                C[s] = kCostInner * C[s] + best;
                P[s] = (uint8_t)(bj == 0xffffffffu ? 0u
                                                   : low | deposit_bits(ncand - 1u - bj, rest));
            }
            wave_lds_step();
        }
        // the full set: its 2^(nl-1) - 1 splits over the lanes (low = leaf 0)
        const uint32_t rest = full ^ 1u, nsplit = (1u << (nl - 1)) - 1u;
        float c = INFINITY;
        if (lane < nsplit) {
            const uint32_t sub = (nsplit - 1u - lane) << 1;
            c = C[1u | sub] + C[rest ^ sub];
        }
        const float cmin = wave_min(c);
        // best_split's choice: its order is the lanes' order, so the first
        // lane at the least cost, none when no cost is below infinity
        const uint64_t at = __ballot(lane < nsplit && c == cmin && cmin < INFINITY);
        const uint32_t pfull = at ? 1u | ((nsplit - 1u - (uint32_t)__builtin_ctzll(at)) << 1) : 0u;
        const float cf = kCostInner * C[full] + cmin;  // C[full]: its surface area
        wave_lds_step();  // every lane has read C[full] before lane 0 replaces it
        if (lane == 0) {
            C[full] = cf;
            P[full] = (uint8_t)pfull;
        }
        wave_lds_step();
        if (lane == 0 && cf < cx) {
            // plan the rebuild first (no stores): breadth-first over the subsets
            int32_t qn[kTreeletLeaves - 1], ql[kTreeletLeaves - 1], qr[kTreeletLeaves - 1];
            uint32_t qs[kTreeletLeaves - 1];
            qn[0] = x;
            qs[0] = full;
            int nq = 1, next = 1;
            bool ok = true;
            for (int i = 0; i < nq && ok; ++i) {
                const uint32_t p = P[qs[i]];
                if (p == 0) {
                    ok = false;
                    break;
                }
                const uint32_t part[2] = {p, qs[i] ^ p};
                int32_t ch[2];
                for (int j = 0; j < 2; ++j) {
                    if (!(part[j] & (part[j] - 1))) {
                        ch[j] = lf[__ffs(part[j]) - 1];
                    } else {
                        ch[j] = in[next++];
                        qn[nq] = ch[j];
                        qs[nq] = part[j];
                        ++nq;
                    }
                }
                ql[i] = ch[0];
                qr[i] = ch[1];
            }
            if (ok) {
                uint32_t elo[kTreeletLeaves], ehi[kTreeletLeaves], ec[kTreeletLeaves];
#pragma unroll
                for (int q = 0; q < kTreeletLeaves; ++q) {
                    elo[q] = __builtin_amdgcn_readlane(mlo, q);
                    ehi[q] = __builtin_amdgcn_readlane(mhi, q);
                    ec[q] = __builtin_amdgcn_readlane(mc, q);
                }
                for (int i = 0; i < nq; ++i) {
                    A.left[qn[i]] = ql[i];
                    A.right[qn[i]] = qr[i];
                    A.parent[ql[i]] = qn[i];
                    A.parent[qr[i]] = qn[i];
                }
                for (int i = 1; i < nq; ++i) {
                    const int32_t y = qn[i];
                    float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
                    uint32_t lo = 0xffffffffu, hi = 0, c = 0;
#pragma unroll
                    for (int q = 0; q < kTreeletLeaves; ++q) {
                        if (!((qs[i] >> q) & 1u)) continue;
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            b[a] = fminf(b[a], lb[q][a]);
                            b[3 + a] = fmaxf(b[3 + a], lb[q][3 + a]);
                        }
                        lo = min(lo, elo[q]);
                        hi = max(hi, ehi[q]);
                        c += ec[q];
                    }
                    float* by = A.box + 6 * (size_t)y;
#pragma unroll
                    for (int a = 0; a < 6; ++a) by[a] = b[a];
                    A.emin[y] = lo;
                    A.emax[y] = hi;
                    A.cnt[y] = c;
                    A.cost[y] = C[qs[i]];
                }
                cx = cf;
            }
        }
    }
    if (lane == 0) A.cost[x] = cx;
}

// Windows of a restructured tree.  Restructuring regroups subtrees, so a
// node's triangles stop being a range of sorted positions, and with windows
// over positions most nodes of a terrain came to straddle window boundaries
// after a few rounds (most of a round's nodes then went to the agent-scope
// pass).  Each pass therefore first ranks the leaves in the current tree's
// order (left subtree before right): every subtree is a range of ranks
// [emin, emax], and the windows are over ranks.  rank_start(x) = the sum,
// over the ancestors a of x reached from a's right child, of the triangles
// in a's left child: pointer jumping over (ancestor, partial sum) pairs,
// each a single 64-bit word, so any pair read is a consistent one and every
// thread's chain of reads ends at the root whatever it sees of the others'
// progress.

// The triangle counts of the radix tree (before the first round).
__global__ __launch_bounds__(256) void k_tl_count(RadixArrays A) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < A.m) A.cnt[i] = A.rl[i] - A.rf[i] + 1;
}
__global__ void k_tl_zero(RadixArrays A) {
    if (threadIdx.x < 2) A.top_count[threadIdx.x] = 0;
}
__device__ __forceinline__ uint64_t rank_pair(int32_t anc, uint32_t sum) {
    return ((uint64_t)(uint32_t)anc << 32) | sum;
}
__global__ __launch_bounds__(256) void k_tl_rank_init(RadixArrays A) {
    const uint32_t x = blockIdx.x * 256 + threadIdx.x;
    if (x >= A.m) return;
    const int32_t p = A.parent[x];
    const uint32_t off = (p >= 0 && A.right[p] == (int32_t)x) ? ncount(A, A.left[p]) : 0u;
    A.jump[x] = rank_pair(p, off);
}
// Ranks of the internal nodes, then arrival counters cleared and the nodes
// whose rank ranges cross a window boundary listed (top_count[0]; zeroed
// with the start count top_count[1] by k_tl_zero before).
__global__ __launch_bounds__(256) void k_tl_rank_jump(RadixArrays A) {
    const uint32_t x = blockIdx.x * 256 + threadIdx.x;
    if (x >= A.m) return;
    uint64_t v = A.jump[x];
    for (uint32_t steps = 0; (int32_t)(v >> 32) >= 0 && steps <= A.m; ++steps) {
        const uint64_t w = __hip_atomic_load(A.jump + (uint32_t)(v >> 32), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        v = (w & 0xffffffff00000000ull) | (uint32_t)((uint32_t)v + (uint32_t)w);
        __hip_atomic_store(A.jump + x, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const uint32_t r = (uint32_t)v, last = r + A.cnt[x] - 1;
    A.emin[x] = r;
    A.emax[x] = last;
    A.done[x] = 0;
    if (r >> A.wshift != last >> A.wshift) A.top[atomicAdd(A.top_count, 1u)] = x;
}
// Ranks of the sorted positions, and the position of each rank.
__global__ __launch_bounds__(256) void k_tl_rank_leaves(RadixArrays A) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= A.n) return;
    const int32_t u = (int32_t)(A.m + p), x = A.parent[u];
    const uint32_t r = A.emin[x] + (A.right[x] == u ? ncount(A, A.left[x]) : 0u);
    A.rank[p] = r;
    if (r < A.n) A.order[r] = p;  // always: the ranks are a permutation of [0, n)
}

// What a bottom-up pass over the restructured tree does at a node: a
// treelet step (the rounds), or the Box4 candidates and sizes (once, after
// the last round: treelet steps change no box they do not write, so the
// final boxes are there).
enum TlMode { kTlTreelet, kTlBox4 };

// Bottom-up climbs of one block: lane climbs start at `start` (or none: -1);
// at every step each lane checks its node (window test for pass 1) and counts
// its arrival; the second arrival at a node owns it.  kTlTreelet: the lanes
// form their owned nodes' treelets, then the wave runs the treelet steps one
// after another; kTlBox4: each lane finishes its own node.  Arrival scope:
// workgroup for the windows (every node inside a window is reached only from
// its block), agent for the nodes crossing windows (reached from any block).
template <bool WINDOW, int MODE>
__device__ void tl_climbs(const RadixArrays& A, int32_t x, uint32_t lo, uint32_t hi,
                          TlShared* sh, const uint8_t* subsets) {
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t steps = 0;; ++steps) {
        bool own = false;
        if (x >= 0 && steps <= A.m) {
            if constexpr (WINDOW) {
                if (A.emin[x] >= lo && A.emax[x] <= hi)
                    own = __hip_atomic_fetch_add(A.done + x, 1u, __ATOMIC_ACQ_REL,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP) != 0;
            } else {
                // the climb's first node (steps 0) is a start: both children
                // finished before this pass
                own = steps == 0 || __hip_atomic_fetch_add(A.done + x, 1u, __ATOMIC_ACQ_REL,
                                                            __HIP_MEMORY_SCOPE_AGENT) != 0;
            }
        }
        if (!own) x = -1;
        uint64_t m = __ballot(own);
        if (m == 0) break;
        // the wave reads what other lanes' (and, for agent scope, other
        // blocks') steps wrote before their arrivals
        if constexpr (WINDOW) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        else __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if constexpr (MODE == kTlBox4) {
            if (own) finish_box4(A, x);
        } else {
            TlForm f;
            f.nl = 0;
            float* rows = sh->row;
            if (own) treelet_form(A, x, f, rows + lane * kRowFloats);
            while (m) {
                const int l = __builtin_ctzll(m);
                m &= m - 1;
                const int nl = __builtin_amdgcn_readlane(f.nl, l);
                if (nl == 0) continue;
                int32_t lf[kTreeletLeaves], in[kTreeletLeaves - 1];
#pragma unroll
                for (int q = 0; q < kTreeletLeaves; ++q)
                    lf[q] = __builtin_amdgcn_readlane(f.lf[q], l);
#pragma unroll
                for (int q = 0; q < kTreeletLeaves - 1; ++q)
                    in[q] = __builtin_amdgcn_readlane(f.in[q], l);
                const float cx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(f.cx), l));
                wave_lds_step();  // the rows written above
                treelet_step(A, __builtin_amdgcn_readlane(x, l), nl, cx, lf, in,
                             rows + l * kRowFloats, sh, subsets);
            }
        }
        if constexpr (WINDOW) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        else __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (own) x = A.parent[x];
        (void)lane;
    }
}

constexpr uint32_t kTreeletWaves = kTreeletBlock / 64;
static_assert(kTreeletWaves * sizeof(TlShared) + kTreeletSets <= 53760,
              "treelet blocks: 3 to a CU (LDS)");

// One restructuring round (or the Box4 pass), pass 1: as k_rt_bottom_up over
// windows of 1 << wshift leaf ranks, with a node inside the window when all
// ranks below it are.
template <int MODE>
__global__ __launch_bounds__(kTreeletBlock, 3) void k_tl_window(RadixArrays A) {
    __shared__ TlShared sh[MODE == kTlTreelet ? kTreeletWaves : 1];
    __shared__ uint8_t subsets[kTreeletSets];
    if (MODE == kTlTreelet && threadIdx.x < kTreeletSets)
        subsets[threadIdx.x] = kSubsetsBySize[threadIdx.x];
    __syncthreads();
    TlShared* mine = sh + (MODE == kTlTreelet ? threadIdx.x >> 6 : 0);
    const uint32_t lo = blockIdx.x << A.wshift;
    const uint32_t hi = min(lo + (1u << A.wshift), A.n) - 1;
    for (uint32_t base = lo; base <= hi; base += kTreeletBlock) {
        const uint32_t r = base + threadIdx.x;
        tl_climbs<true, MODE>(A, r <= hi ? A.parent[A.m + A.order[r]] : -1, lo, hi, mine,
                              subsets);
    }
}

// Pass 2 set-up: the crossing nodes whose children both lie inside windows
// (finished by pass 1) start climbs; one whose single such child counts as
// an arrival up front.
__global__ __launch_bounds__(256) void k_tl_top_prep(RadixArrays A) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= *A.top_count) return;
    const int32_t x = (int32_t)A.top[i];
    const int ready =
        (ext_straddles(A, A.left[x]) ? 0 : 1) + (ext_straddles(A, A.right[x]) ? 0 : 1);
    if (ready == 1)
        A.done[x] = 1;
    else if (ready == 2)
        A.starts[atomicAdd(A.top_count + 1, 1u)] = (uint32_t)x;
}

// Pass 2: the crossing nodes, climbed from the starts (agent scope
// arrivals).  kTlTreelet: one climb per wave (its steps are the wave's
// work, and a wave's climbs run one after another); kTlBox4: one per lane.
template <int MODE>
__global__ __launch_bounds__(kTreeletBlock, 3) void k_tl_top(RadixArrays A) {
    __shared__ TlShared sh[MODE == kTlTreelet ? kTreeletWaves : 1];
    __shared__ uint8_t subsets[kTreeletSets];
    const uint32_t ns = A.top_count[1];
    if (blockIdx.x * kTreeletWaves >= ns) return;  // no climb for this block (uniform)
    if (MODE == kTlTreelet && threadIdx.x < kTreeletSets)
        subsets[threadIdx.x] = kSubsetsBySize[threadIdx.x];
    __syncthreads();
    if constexpr (MODE == kTlTreelet) {
        const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
        for (uint32_t j = blockIdx.x * kTreeletWaves + wave; j < ns;
             j += gridDim.x * kTreeletWaves)
            tl_climbs<false, MODE>(A, lane == 0 ? (int32_t)A.starts[j] : -1, 0, 0, sh + wave,
                                   subsets);
    } else {
        for (uint32_t base = blockIdx.x * kTreeletBlock; base < ns;
             base += gridDim.x * kTreeletBlock) {
            const uint32_t j = base + threadIdx.x;
            tl_climbs<false, MODE>(A, j < ns ? (int32_t)A.starts[j] : -1, 0, 0, sh, subsets);
        }
    }
}

// Internal node x with more than maxp triangles: search down from the root
// through the candidates; if x is a Box4 node, its pre-order index is the
// parent's index + 1 + the sizes of the inner candidates before it.
// WALK (a restructured tree, whose subtrees are not ranges): the candidate on
// the way to x is the one met first when walking up from x through parent
// links; none before reaching cur means x was expanded into cur's node.
template <bool WALK>
__global__ __launch_bounds__(256) void k_rt_emit(RadixArrays A, rt_node4* __restrict__ nodes,
                                                 rt_build_info* info) {
    const uint32_t x = blockIdx.x * 256 + threadIdx.x;
    if (x >= A.m) return;
    const uint32_t xf = A.rf[x], xl = A.rl[x];
    if (xl - xf + 1 <= A.maxp) return;
    int32_t cur = 0;
    uint32_t idx = 0, level = 1;
    while (cur != (int32_t)x) {
        const int4 cv = A.cand[cur];
        const int32_t cs[4] = {cv.x, cv.y, cv.z, cv.w};
        uint32_t acc = idx + 1;
        int32_t next = -1;
        int on_path = -1;
        if constexpr (WALK) {
            int32_t u = (int32_t)x;
            for (uint32_t steps = 0; u >= 0 && u != cur && on_path < 0 && steps <= A.m;
                 ++steps, u = A.parent[u])
                for (int q = 0; q < 4; ++q)
                    if (cs[q] == u) on_path = q;
        }
        for (int q = 0; q < 4 && cs[q] >= 0; ++q) {
            const int32_t c = cs[q];
            const bool inner = ninner(A, c);
            const bool hit = WALK ? q == on_path : (nfirst(A, c) <= xf && xl <= nlast(A, c));
            if (hit) {
                if (inner) next = c;
                break;
            }
            if (inner) acc += A.size4[c];
        }
        if (next < 0) return;  // x was expanded into its parent's Box4 node
        if (WALK && level > A.m) return;  // deeper than any tree of m inner nodes
        cur = next;
        idx = acc;
        ++level;
    }
    const int4 cv = A.cand[x];
    const int32_t cs[4] = {cv.x, cv.y, cv.z, cv.w};
    NodeReg nr;
    node_clear(nr);
    uint32_t acc = idx + 1;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int32_t c = cs[q];
        if (c < 0) continue;
        float bb[6];
#pragma unroll
        for (int a = 0; a < 6; ++a) bb[a] = A.box[6 * (size_t)c + a];
        if (ninner(A, c)) {
            node_set(nr, q, bb, (int32_t)acc, -1);
            acc += A.size4[c];
        } else {
            const uint32_t f = nfirst(A, c);
            node_set(nr, q, bb, (int32_t)~f, (int32_t)(nlast(A, c) - f + 1));
        }
    }
    node_store(nodes + idx, nr);
    atomicMax(&info->depth, level);
    if (x == 0) info->node_count = A.size4[0];
}

// n <= maxp: the whole mesh is one leaf range; a root with one slot
// (software_bvh.hpp:205-219).
__global__ void k_rt_single(const float* __restrict__ box, uint32_t n, uint32_t m,
                            rt_node4* __restrict__ nodes, rt_build_info* info) {
    if (threadIdx.x != 0) return;
    float bb[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (uint32_t p = 0; p < n; ++p)
        for (int a = 0; a < 3; ++a) {
            bb[a] = fminf(bb[a], box[6 * (size_t)(m + p) + a]);
            bb[3 + a] = fmaxf(bb[3 + a], box[6 * (size_t)(m + p) + 3 + a]);
        }
    NodeReg nr;
    node_clear(nr);
    node_set(nr, 0, bb, (int32_t)~0u, (int32_t)n);
    node_store(nodes, nr);
    info->node_count = 1;
    info->depth = 1;
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

// Level table of the fused path (t.levels <= 6 for n <= kSmallMax); unused
// entries stay zero.
SmallTree small_tree(const TreeLayout& t) {
    SmallTree a = {};
    a.levels = t.levels;
    uint32_t pad = 0;
    for (uint32_t i = 0; i < t.levels && i < (uint32_t)kSmallLevels; ++i) {
        a.count[i] = t.counts[i];
        a.off[i] = t.offsets[i];
        a.pad[i] = pad;
        pad += (t.counts[i] + 3u) & ~3u;
    }
    return a;
}
static_assert(kSmallMax == 4096 && kBoxPitch == 4 + 4 + 16 + 64 + 256 + 1024,
              "LDS box pitch covers the padded levels of a 4096-triangle tree");

struct ScratchPlan {
    size_t bounds = 0, keys0 = 0, vals0 = 0, keys1 = 0, vals1 = 0, hist = 0, totals = 0,
           boxes = 0, rbox = 0, rf = 0, rl = 0, left = 0, right = 0, parent = 0, done = 0,
           top = 0, top_count = 0, starts = 0, cand = 0, size4 = 0, cost = 0, emin = 0,
           emax = 0, cnt = 0, jump = 0, rank = 0, order = 0, total = 0;
    uint32_t tiles = 0;
};

// Treelet windows: the window pass runs one block per window with 3 blocks
// to a CU, its waves taking their nodes' treelets one after another, so
// windows small enough for every CU to hold blocks, and for blocks to
// outnumber the resident slots (balance); larger windows leave fewer nodes
// crossing window boundaries to the agent-scope pass.
uint32_t treelet_window_shift(uint32_t n) {
    uint32_t sh = 12;  // 4096
    while (sh > 9 && ((size_t)n >> sh) < 1536) --sh;
    return sh;
}

// Scratch of one build: sort buffers, then the implicit level boxes or the
// radix-tree arrays.
ScratchPlan plan_scratch(uint32_t n, uint32_t algo, uint32_t rounds) {
    ScratchPlan p;
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    size_t off = 0;
    const size_t m = n ? n - 1 : 0, u = m + n;
    p.tiles = ceil_div(n, kTile);
    p.bounds = off; off += al(8 * 4);
    p.keys0 = off; off += al((size_t)n * 4);
    p.vals0 = off; off += al((size_t)n * 4);
    p.keys1 = off; off += al((size_t)n * 4);
    p.vals1 = off; off += al((size_t)n * 4);
    p.hist = off; off += al((size_t)kMaxDigits * p.tiles * 4);
    p.totals = off; off += al(kMaxDigits * 4);
    if (algo == RT_BUILD_IMPLICIT) {
        p.boxes = off; off += al((size_t)tree_layout(n).total_boxes * 24);
    } else {
        p.rbox = off; off += al(u * 24);
        p.rf = off; off += al(m * 4);
        p.rl = off; off += al(m * 4);
        p.left = off; off += al(m * 4);
        p.right = off; off += al(m * 4);
        p.parent = off; off += al(u * 4);
        p.done = off; off += al(m * 4);
        p.top = off; off += al(m * 4);
        p.top_count = off; off += al(4);
        // climb starts: one per window boundary in the radix tree, but a
        // restructured tree's subtrees are no longer ranges: room for every node
        p.starts = off; off += al(std::max(m, (size_t)n / kWindow + 1) * 4);
        p.cand = off; off += al(m * 16);
        p.size4 = off; off += al(m * 4);
        p.cost = off; off += al(m * 4);
        p.emin = off; off += al(m * 4);
        p.emax = off; off += al(m * 4);
        if (rounds) {
            p.cnt = off; off += al(m * 4);
            p.jump = off; off += al(m * 8);
            p.rank = off; off += al((size_t)n * 4);
            p.order = off; off += al((size_t)n * 4);
        }
    }
    p.total = off;
    return p;
}

uint32_t resolve_algo(const rt_build_cfg* cfg, uint32_t n) {
    const uint32_t a = cfg ? cfg->algorithm : RT_BUILD_AUTO;
    if (a == RT_BUILD_AUTO) return n <= kSmallMax ? RT_BUILD_IMPLICIT : RT_BUILD_RADIX;
    return a;
}

// Digit width of the multi-launch radix sort (the sort's result does not
// depend on it): 0 = 8 bits; 4, 6, 8 or 10 bits.
uint32_t resolve_bits(const rt_build_cfg* cfg) {
    const uint32_t b = cfg ? cfg->sort_bits : 0u;
    if (b == 0) return 8u;
    return (b == 4 || b == 6 || b == 8 || b == 10) ? b : 0u;
}

uint32_t resolve_maxp(const rt_build_cfg* cfg) {
    return (cfg && cfg->max_leaf_prims) ? cfg->max_leaf_prims : 4u;
}

uint32_t max_nodes(uint32_t n, uint32_t algo) {
    if (algo == RT_BUILD_IMPLICIT) return tree_layout(n).total_boxes;
    return n > 1 ? n - 1 : 1;
}

__global__ void k_set_info(rt_build_info* info, uint32_t nodes, uint32_t depth) {
    if (threadIdx.x == 0) {
        info->node_count = nodes;
        info->depth = depth;
    }
}

template <int BITS>
rt_status sort_passes(hipStream_t s, uint32_t n, const ScratchPlan& p, uint32_t* hist,
                      uint32_t* totals, uint32_t** ki, uint32_t** vi, uint32_t** ko,
                      uint32_t** vo) {
    constexpr uint32_t D = 1u << BITS;
    for (uint32_t shift = 0; shift < 30; shift += BITS) {  // codes have 30 bits
        hipLaunchKernelGGL(k_hist<BITS>, dim3(p.tiles), dim3(kSortBlock), 0, s, *ki, n, shift,
                           p.tiles, hist);
        hipLaunchKernelGGL(k_digit_scan, dim3(D), dim3(1024), 0, s, hist, p.tiles, totals);
        hipLaunchKernelGGL(k_scatter<BITS>, dim3(p.tiles), dim3(kSortBlock), 0, s, *ki, *vi, *ko,
                           *vo, n, shift, p.tiles, hist, totals);
        const rt_status st = launch_status("k_scatter");
        if (st != RT_OK) return st;
        std::swap(*ki, *ko);
        std::swap(*vi, *vo);
    }
    return RT_OK;
}

// Bounds -> Morton -> LSD radix sort of BITS-bit digits (ceil(30 / BITS)
// stable passes); *keys / *vals point at the sorted codes / ids (in either
// pair of sort buffers).
rt_status morton_sort(hipStream_t s, const float* pos, const uint32_t* idx, uint32_t n,
                      uint32_t bits, const ScratchPlan& p, char* base, uint32_t** keys,
                      uint32_t** vals) {
    uint32_t* ob = reinterpret_cast<uint32_t*>(base + p.bounds);
    uint32_t* k0 = reinterpret_cast<uint32_t*>(base + p.keys0);
    uint32_t* v0 = reinterpret_cast<uint32_t*>(base + p.vals0);
    uint32_t* k1 = reinterpret_cast<uint32_t*>(base + p.keys1);
    uint32_t* v1 = reinterpret_cast<uint32_t*>(base + p.vals1);
    uint32_t* hist = reinterpret_cast<uint32_t*>(base + p.hist);
    uint32_t* totals = reinterpret_cast<uint32_t*>(base + p.totals);
    hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, s, ob);
    const uint32_t bgrid = std::min<uint32_t>(ceil_div(n, 256), 2048);
    hipLaunchKernelGGL(k_bounds, dim3(bgrid), dim3(256), 0, s, pos, idx, n, ob);
    hipLaunchKernelGGL(k_morton, dim3(ceil_div(n, 256)), dim3(256), 0, s, pos, idx, n, ob, k0, v0);
    rt_status st = launch_status("k_morton");
    if (st != RT_OK) return st;
    uint32_t *ki = k0, *vi = v0, *ko = k1, *vo = v1;
    switch (bits) {
        case 4: st = sort_passes<4>(s, n, p, hist, totals, &ki, &vi, &ko, &vo); break;
        case 6: st = sort_passes<6>(s, n, p, hist, totals, &ki, &vi, &ko, &vo); break;
        case 10: st = sort_passes<10>(s, n, p, hist, totals, &ki, &vi, &ko, &vo); break;
        default: st = sort_passes<8>(s, n, p, hist, totals, &ki, &vi, &ko, &vo); break;
    }
    *keys = ki;
    *vals = vi;
    return st;
}

rt_status build_radix(hipStream_t s, const float* pos, const uint32_t* idx, uint32_t n,
                      uint32_t maxp, uint32_t rounds, const ScratchPlan& p, char* base,
                      const uint32_t* keys,
                      const uint32_t* vals, rt_node4* nodes, rt_tri64* tris,
                      rt_build_info* info) {
    RadixArrays A;
    A.keys = keys;
    A.box = reinterpret_cast<float*>(base + p.rbox);
    A.rf = reinterpret_cast<uint32_t*>(base + p.rf);
    A.rl = reinterpret_cast<uint32_t*>(base + p.rl);
    A.left = reinterpret_cast<int32_t*>(base + p.left);
    A.right = reinterpret_cast<int32_t*>(base + p.right);
    A.parent = reinterpret_cast<int32_t*>(base + p.parent);
    A.done = reinterpret_cast<uint32_t*>(base + p.done);
    A.top = reinterpret_cast<uint32_t*>(base + p.top);
    A.top_count = reinterpret_cast<uint32_t*>(base + p.top_count);
    A.starts = reinterpret_cast<uint32_t*>(base + p.starts);
    A.cand = reinterpret_cast<int4*>(base + p.cand);
    A.size4 = reinterpret_cast<uint32_t*>(base + p.size4);
    A.cost = reinterpret_cast<float*>(base + p.cost);
    A.emin = reinterpret_cast<uint32_t*>(base + p.emin);
    A.emax = reinterpret_cast<uint32_t*>(base + p.emax);
    A.cnt = rounds ? reinterpret_cast<uint32_t*>(base + p.cnt) : nullptr;
    A.jump = rounds ? reinterpret_cast<uint64_t*>(base + p.jump) : nullptr;
    A.rank = rounds ? reinterpret_cast<uint32_t*>(base + p.rank) : nullptr;
    A.order = rounds ? reinterpret_cast<uint32_t*>(base + p.order) : nullptr;
    A.n = n;
    A.m = n - 1;
    A.maxp = maxp;
    A.wshift = treelet_window_shift(n);
    hipLaunchKernelGGL(k_rt_leaves, dim3(ceil_div(n, 256)), dim3(256), 0, s, pos, idx, n, vals,
                       tris, A.box, A.m, A.top_count);
    if (n <= maxp) {
        hipLaunchKernelGGL(k_rt_single, dim3(1), dim3(64), 0, s, A.box, n, A.m, nodes, info);
        return launch_status("k_rt_single");
    }
    hipLaunchKernelGGL(k_rt_radix, dim3(ceil_div(A.m, 256)), dim3(256), 0, s, A, info);
    hipLaunchKernelGGL(k_rt_bottom_up, dim3(ceil_div(n, kWindow)), dim3(1024), 0, s, A);
    hipLaunchKernelGGL(k_rt_top, dim3(1), dim3(1024), 0, s, A);
    // restructuring rounds, then the Box4 candidates of the final tree; each
    // pass ranks the leaves of the current tree first
    if (rounds) hipLaunchKernelGGL(k_tl_count, dim3(ceil_div(A.m, 256)), dim3(256), 0, s, A);
    for (uint32_t r = 0; r <= rounds && rounds; ++r) {
        hipLaunchKernelGGL(k_tl_zero, dim3(1), dim3(64), 0, s, A);
        hipLaunchKernelGGL(k_tl_rank_init, dim3(ceil_div(A.m, 256)), dim3(256), 0, s, A);
        hipLaunchKernelGGL(k_tl_rank_jump, dim3(ceil_div(A.m, 256)), dim3(256), 0, s, A);
        hipLaunchKernelGGL(k_tl_rank_leaves, dim3(ceil_div(n, 256)), dim3(256), 0, s, A);
        if (r < rounds)
            hipLaunchKernelGGL(k_tl_window<kTlTreelet>, dim3(ceil_div(n, 1u << A.wshift)),
                               dim3(kTreeletBlock), 0, s, A);
        else
            hipLaunchKernelGGL(k_tl_window<kTlBox4>, dim3(ceil_div(n, 1u << A.wshift)),
                               dim3(kTreeletBlock), 0, s, A);
        hipLaunchKernelGGL(k_tl_top_prep, dim3(ceil_div(A.m, 256)), dim3(256), 0, s, A);
        if (r < rounds)
            hipLaunchKernelGGL(k_tl_top<kTlTreelet>, dim3(kTreeletTopBlocks), dim3(kTreeletBlock),
                               0, s, A);
        else
            hipLaunchKernelGGL(k_tl_top<kTlBox4>, dim3(kTreeletTopBlocks), dim3(kTreeletBlock), 0,
                               s, A);
    }
    if (rounds)
        hipLaunchKernelGGL(k_rt_emit<true>, dim3(ceil_div(A.m, 256)), dim3(256), 0, s, A, nodes,
                           info);
    else
        hipLaunchKernelGGL(k_rt_emit<false>, dim3(ceil_div(A.m, 256)), dim3(256), 0, s, A, nodes,
                           info);
    return launch_status("k_rt_emit");
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

extern "C" rt_status rt_build_requirements(uint32_t n, const rt_build_cfg* cfg,
                                           uint32_t* node_capacity, size_t* scratch_bytes) {
    const uint32_t algo = resolve_algo(cfg, n);
    if (algo != RT_BUILD_IMPLICIT && algo != RT_BUILD_RADIX) return RT_ERR_ARG;
    if (resolve_bits(cfg) == 0) return RT_ERR_ARG;
    if (node_capacity) *node_capacity = max_nodes(n, algo);
    if (scratch_bytes) {
        const bool fused = algo == RT_BUILD_IMPLICIT && n <= kSmallMax;
        *scratch_bytes = fused ? 0 : plan_scratch(n, algo, cfg ? cfg->treelet_rounds : 0u).total;
    }
    return RT_OK;
}

extern "C" rt_status rt_gpu_build(rt_stream stream, const float* d_pos, const uint32_t* d_idx,
                                  uint32_t n, const rt_build_cfg* cfg, rt_node4* d_nodes,
                                  uint32_t node_capacity, rt_tri64* d_tris, rt_build_info* d_info,
                                  uint32_t* d_sorted_ids, uint32_t* d_codes, void* d_scratch,
                                  size_t scratch_bytes) {
    if (!d_info) return RT_ERR_ARG;
    hipStream_t s = as_stream(stream);
    if (n == 0) {
        hipLaunchKernelGGL(k_set_info, dim3(1), dim3(64), 0, s, d_info, 0u, 0u);
        return launch_status("k_set_info");
    }
    if (!d_pos || !d_nodes || !d_tris) return RT_ERR_ARG;
    if (n > (1u << 30)) return RT_ERR_CAPACITY;
    const uint32_t algo = resolve_algo(cfg, n);
    const uint32_t maxp = resolve_maxp(cfg);
    const uint32_t bits = resolve_bits(cfg);
    const uint32_t rounds = cfg ? cfg->treelet_rounds : 0u;
    if (algo != RT_BUILD_IMPLICIT && algo != RT_BUILD_RADIX) return RT_ERR_ARG;
    if (bits == 0) return RT_ERR_ARG;
    if (maxp < 1 || maxp > 16) return RT_ERR_ARG;
    if (rounds > kMaxTreeletRounds || (rounds && algo != RT_BUILD_RADIX)) return RT_ERR_ARG;
    if (node_capacity < max_nodes(n, algo)) return RT_ERR_CAPACITY;
    const TreeLayout t = tree_layout(n);
    if (algo == RT_BUILD_IMPLICIT && n <= kSmallMax) {
        hipLaunchKernelGGL(k_build_small, dim3(ceil_div(n, kPartRecs)), dim3(kSmallBlock), 0, s,
                           d_pos, d_idx, n, d_nodes, d_tris, d_sorted_ids, d_codes,
                           small_tree(t), d_info);
        return launch_status("k_build_small");
    }
    const ScratchPlan p = plan_scratch(n, algo, rounds);
    if (!d_scratch || scratch_bytes < p.total) return RT_ERR_CAPACITY;
    char* base = static_cast<char*>(d_scratch);
    uint32_t *keys = nullptr, *vals = nullptr;
    rt_status st = morton_sort(s, d_pos, d_idx, n, bits, p, base, &keys, &vals);
    if (st != RT_OK) return st;
    if (d_sorted_ids)
        RT_TRY_HIP(hipMemcpyAsync(d_sorted_ids, vals, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
    if (d_codes)
        RT_TRY_HIP(hipMemcpyAsync(d_codes, keys, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
    if (algo == RT_BUILD_RADIX)
        return build_radix(s, d_pos, d_idx, n, maxp, rounds, p, base, keys, vals, d_nodes, d_tris,
                           d_info);
    float* gboxes = reinterpret_cast<float*>(base + p.boxes);
    st = emit_levels(s, d_pos, d_idx, n, vals, d_nodes, d_tris, gboxes, t, false);
    if (st != RT_OK) return st;
    hipLaunchKernelGGL(k_set_info, dim3(1), dim3(64), 0, s, d_info, t.total_boxes, t.levels);
    return launch_status("k_set_info");
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
        hipLaunchKernelGGL(k_refit_small, dim3(ceil_div(n, kPartRecs)), dim3(kSmallBlock), 0, s,
                           d_pos, d_idx, n, d_nodes, d_tris, small_tree(t), nullptr);
        return launch_status("k_refit_small");
    }
    const ScratchPlan p = plan_scratch(n, RT_BUILD_IMPLICIT, 0);
    if (!d_scratch || scratch_bytes < p.total) return RT_ERR_CAPACITY;
    float* gboxes = reinterpret_cast<float*>(static_cast<char*>(d_scratch) + p.boxes);
    return emit_levels(s, d_pos, d_idx, n, nullptr, d_nodes, d_tris, gboxes, t, true);
}

extern "C" rt_status rt_gpu_refit_bvh4(rt_stream stream, const float* d_pos,
                                       const uint32_t* d_idx, rt_node4* d_nodes,
                                       uint32_t node_count, rt_tri64* d_tris, uint32_t depth,
                                       void* d_scratch, size_t scratch_bytes) {
    if (node_count == 0) return RT_OK;
    if (!d_pos || !d_nodes || !d_tris) return RT_ERR_ARG;
    if (depth > kMaxDepth) return RT_ERR_ARG;
    if (!d_scratch || scratch_bytes < (size_t)node_count * 4) return RT_ERR_CAPACITY;
    hipStream_t s = as_stream(stream);
    uint32_t* done = static_cast<uint32_t*>(d_scratch);
    const dim3 grid(ceil_div(node_count, 256)), block(256);
    hipLaunchKernelGGL(k_refit_leaf_slots, grid, block, 0, s, d_pos, d_idx, d_nodes, node_count,
                       d_tris, done);
    const uint32_t rounds = depth ? depth : kMaxDepth;
    for (uint32_t r = 0; r < rounds; ++r)
        hipLaunchKernelGGL(k_refit_round, grid, block, 0, s, d_nodes, node_count, done, r);
    return launch_status("k_refit_round");
}

extern "C" rt_status rt_gpu_build_batch_table_bytes(uint32_t count, size_t* bytes) {
    if (!bytes) return RT_ERR_ARG;
    *bytes = (size_t)count * sizeof(BatchBuild);
    return RT_OK;
}

extern "C" rt_status rt_gpu_build_batch_plan(const rt_build_item* items, uint32_t count,
                                             void* table, size_t table_bytes) {
    if ((count && (!items || !table)) || table_bytes < (size_t)count * sizeof(BatchBuild))
        return RT_ERR_ARG;
    BatchBuild* out = static_cast<BatchBuild*>(table);
    for (uint32_t k = 0; k < count; ++k) {
        const rt_build_item& it = items[k];
        if (it.n == 0 || it.n > RT_BATCH_BUILD_MAX) return RT_ERR_CAPACITY;
        if (!it.pos || !it.nodes || !it.tris || !it.info) return RT_ERR_ARG;
        const TreeLayout t = tree_layout(it.n);
        if (it.node_capacity < t.total_boxes) return RT_ERR_CAPACITY;
        BatchBuild e = {};
        e.item = it;
        if (it.n <= kSmallMax) {
            e.tree = small_tree(t);
        } else {
            // the multi-launch implicit build's scratch layout
            const ScratchPlan p = plan_scratch(it.n, RT_BUILD_IMPLICIT, 0);
            if (!it.scratch || it.scratch_bytes < p.total) return RT_ERR_CAPACITY;
            MedTree& m = e.med;
            m.levels = t.levels;
            m.total_boxes = t.total_boxes;
            for (uint32_t l = 0; l < t.levels; ++l) {
                m.counts[l] = t.counts[l];
                m.offsets[l] = t.offsets[l];
            }
            m.keys0 = p.keys0;
            m.vals0 = p.vals0;
            m.keys1 = p.keys1;
            m.vals1 = p.vals1;
            m.boxes = p.boxes;
        }
        out[k] = e;
    }
    return RT_OK;
}

namespace {
// small / medium: whether the batch may hold items of each path (both when
// the host does not know the items).
rt_status launch_build_batch(hipStream_t s, const void* d_table, uint32_t count,
                             uint32_t workgroups, bool small, bool medium) {
    const bool shared = (workgroups & RT_BUILD_BATCH_SHARED) != 0;
    workgroups &= ~RT_BUILD_BATCH_SHARED;
    const uint32_t g = workgroups ? min(workgroups, count) : min(count, 256u);
    const auto* t = static_cast<const BatchBuild*>(d_table);
    if (!shared || small)
        hipLaunchKernelGGL(k_build_batch, dim3(g), dim3(kSmallBlock), 0, s, t, count,
                           shared ? 1u : 0u);
    if (shared && medium)
        hipLaunchKernelGGL(k_build_batch_shared, dim3(count), dim3(kSharedBlock), 0, s, t, count);
    return launch_status("k_build_batch");
}
}  // namespace

extern "C" rt_status rt_gpu_build_batch(rt_stream stream, const void* d_table, uint32_t count,
                                        uint32_t workgroups) {
    if (count == 0) return RT_OK;
    if (!d_table) return RT_ERR_ARG;
    return launch_build_batch(as_stream(stream), d_table, count, workgroups, true, true);
}

namespace {
// Byte layout of a batched-build submit table (staging and device copy
// alike): the planned table, then the optional grid generators.
struct BuildSubmitLayout {
    size_t gen = 0, total = 0;
};
BuildSubmitLayout build_submit_layout(uint32_t count, bool with_gen) {
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    BuildSubmitLayout L;
    size_t off = al((size_t)count * sizeof(BatchBuild));
    if (with_gen) {
        L.gen = off;
        off += al((size_t)count * sizeof(GridGen));
    }
    L.total = off;
    return L;
}
}  // namespace

extern "C" rt_status rt_gpu_build_batch_submit_bytes(uint32_t count, int with_gen,
                                                     size_t* bytes) {
    if (!bytes) return RT_ERR_ARG;
    *bytes = build_submit_layout(count, with_gen != 0).total;
    return RT_OK;
}

extern "C" rt_status rt_gpu_build_batch_submit(rt_stream stream, const rt_build_item* items,
                                               const rt_grid_gen* gen, uint32_t count,
                                               uint32_t workgroups, void* pinned,
                                               size_t pinned_bytes, void* d_table,
                                               size_t table_bytes) {
    if (count == 0) return RT_OK;
    if (!items || !pinned || !d_table) return RT_ERR_ARG;
    if (reinterpret_cast<uintptr_t>(d_table) % 256) return RT_ERR_ARG;
    const BuildSubmitLayout L = build_submit_layout(count, gen != nullptr);
    if (pinned_bytes < L.total || table_bytes < L.total) return RT_ERR_CAPACITY;
    char* h = static_cast<char*>(pinned);
    rt_status st = rt_gpu_build_batch_plan(items, count, h, (size_t)count * sizeof(BatchBuild));
    if (st != RT_OK) return st;
    uint32_t max_quads = 0;
    if (gen) {
        GridGen* hg = reinterpret_cast<GridGen*>(h + L.gen);
        for (uint32_t k = 0; k < count; ++k) {
            if (items[k].idx || items[k].n < 2) return RT_ERR_ARG;
            GridGen g;
            g.xyz = const_cast<float*>(items[k].pos);
            dyn_grid_shape(items[k].n, &g.cols, &g.rows);
            g.t = gen[k].t;
            g.extent = gen[k].extent;
            hg[k] = g;
            max_quads = std::max(max_quads, g.cols * g.rows);
        }
    }
    hipStream_t s = as_stream(stream);
    RT_TRY_HIP(hipMemcpyAsync(d_table, pinned, L.total, hipMemcpyHostToDevice, s));
    char* d = static_cast<char*>(d_table);
    if (gen) {
        st = launch_grid_gen(s, reinterpret_cast<const GridGen*>(d + L.gen), count, max_quads);
        if (st != RT_OK) return st;
    }
    bool small = false, medium = false;
    for (uint32_t k = 0; k < count; ++k) (items[k].n <= kSmallMax ? small : medium) = true;
    return launch_build_batch(s, d, count, workgroups, small, medium);
}

extern "C" rt_status rt_gpu_build_radv(rt_stream stream, const float* d_pos,
                                       const uint32_t* d_idx, uint32_t n, uint32_t order,
                                       void* d_image, uint64_t image_bytes,
                                       uint32_t* d_sorted_ids) {
    if (n == 0 || n > kSmallMax) return RT_ERR_CAPACITY;
    if (!d_pos || !d_image || order > RT_RADV_MORTON) return RT_ERR_ARG;
    if (reinterpret_cast<uintptr_t>(d_image) % 64) return RT_ERR_ARG;
    const TreeLayout t = tree_layout(n);
    if (image_bytes < t.total_as_size) return RT_ERR_CAPACITY;
    hipLaunchKernelGGL(k_build_radv, dim3(ceil_div(n, kPartRecs)), dim3(kSmallBlock), 0,
                       as_stream(stream), d_pos, d_idx, n, static_cast<char*>(d_image),
                       t.total_boxes, d_sorted_ids, small_tree(t), order);
    return launch_status("k_build_radv");
}
