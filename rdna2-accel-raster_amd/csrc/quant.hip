// quant.hip -- rt_gpu_quantize: the RT_NODES_Q8 copy of a Box4 tree.
//
// Each rt_node4 becomes a 64-B rt_qnode4 whose four child boxes are 8-bit
// codes on a per-node, per-axis grid p_a + q * 2^e_a, rounded outward, so
// every decoded box contains the slot box it encodes (the layout of
// compressed wide BVHs, Ylitie, Karras & Laine 2017, "Efficient Incoherent
// Ray Traversal on GPUs Through Compressed Wide BVHs", with four children).
// The traversal of a quantized tree reads half the bytes per node visit;
// the vector memory pipe, not arithmetic, bounds the traversal kernels.
//
// The codes are defined by float comparisons of decoded bounds, so they do
// not depend on how a first guess is found:
//   dec(q)  = p + q * 2^e   (q * 2^e exact: one rounding, the add)
//   e       = least e in [-126, 127] with dec(255) >= max bound of the axis
//   qlo[k]  = greatest q in [0, 255] with dec(q) <= slot k's min
//   qhi[k]  = least q in [0, 255] with dec(q) >= slot k's max
// dec is non-decreasing in q and in e, so each is found by stepping from a
// guess until the defining comparison flips.
#include "rt_internal.h"

namespace rt {
namespace {

constexpr int kQBlock = 256;

__device__ __forceinline__ float pow2(int e) { return __int_as_float((e + 127) << 23); }

__device__ __forceinline__ float qdec(float p, int q, float s) { return p + (float)q * s; }

// e of one axis: least e with dec(255) >= hi (hi the axis's max bound).
__device__ int axis_exponent(float p, float hi) {
    int e = -126;
    const double d = (double)hi - (double)p;
    if (d > 0.0 && d < 3.0e38) e = ilogb(d / 255.0);
    e = e < -126 ? -126 : e > 127 ? 127 : e;
    while (e < 127 && !(qdec(p, 255, pow2(e)) >= hi)) ++e;
    while (e > -126 && qdec(p, 255, pow2(e - 1)) >= hi) --e;
    return e;
}

__device__ int code_lo(float p, float s, float lo) {
    if (!(lo >= p)) return 0;  // lo == p, or a NaN bound: the grid's origin
    const double g = ((double)lo - (double)p) / (double)s;
    int q = g >= 255.0 ? 255 : (int)g;
    while (q > 0 && qdec(p, q, s) > lo) --q;
    while (q < 255 && qdec(p, q + 1, s) <= lo) ++q;
    return q;
}

__device__ int code_hi(float p, float s, float hi) {
    if (!(hi >= p)) return hi < p ? 0 : 255;  // below the origin (empty), or NaN
    const double g = ((double)hi - (double)p) / (double)s;
    int q = g >= 255.0 ? 255 : (int)g;
    while (q < 255 && qdec(p, q, s) < hi) ++q;
    while (q > 0 && qdec(p, q - 1, s) >= hi) --q;
    return q;
}

__global__ __launch_bounds__(kQBlock) void k_quantize(const rt_node4* __restrict__ nodes,
                                                      uint32_t count,
                                                      rt_qnode4* __restrict__ out) {
    const uint32_t i = blockIdx.x * kQBlock + threadIdx.x;
    if (i >= count) return;
    const rt_node4 n = nodes[i];
    const float* lo[3] = {n.minX, n.minY, n.minZ};
    const float* hi[3] = {n.maxX, n.maxY, n.maxZ};
    rt_qnode4 q = {};
    for (int a = 0; a < 3; ++a) {
        float p = INFINITY, m = -INFINITY;
        for (int k = 0; k < 4; ++k)
            if (n.count[k] != 0) {
                p = fminf(p, lo[a][k]);
                m = fmaxf(m, hi[a][k]);
            }
        if (!(p <= m)) {  // no occupied slot (or NaN bounds): an empty grid at 0
            p = 0.0f;
            m = 0.0f;
        }
        const int e = axis_exponent(p, m);
        const float s = pow2(e);
        q.p[a] = p;
        q.e[a] = (int8_t)e;
        for (int k = 0; k < 4; ++k) {
            if (n.count[k] == 0) continue;
            q.qlo[a][k] = (uint8_t)code_lo(p, s, lo[a][k]);
            q.qhi[a][k] = (uint8_t)code_hi(p, s, hi[a][k]);
        }
    }
    for (int k = 0; k < 4; ++k) {
        q.count[k] = (int16_t)n.count[k];
        q.child[k] = n.count[k] != 0 ? n.child[k] : 0;
    }
    out[i] = q;
}

}  // namespace
}  // namespace rt

using namespace rt;

static_assert(sizeof(rt_qnode4) == 64, "rt_qnode4 is one 64-B record");

extern "C" rt_status rt_gpu_quantize(rt_stream stream, const rt_node4* d_nodes,
                                     uint32_t node_count, rt_qnode4* d_qnodes) {
    if (node_count == 0) return RT_OK;
    if (!d_nodes || !d_qnodes) return RT_ERR_ARG;
    hipLaunchKernelGGL(k_quantize, dim3(ceil_div(node_count, kQBlock)), dim3(kQBlock), 0,
                       as_stream(stream), d_nodes, node_count, d_qnodes);
    return launch_status("k_quantize");
}
