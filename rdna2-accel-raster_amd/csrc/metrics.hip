// metrics.hip -- image-quality metrics of two float images on the device:
// the mean squared error, PSNR and 8x8-block SSIM of ImageMetrics::compute
// (common/image_metrics.hpp:16-104), in double precision.
//
// Each thread owns whole 8x8 blocks (grid-stride over the blocks in
// row-major block order): it loads the block's 64 pixels of both images into
// registers, sums them in row-major order for the means, then sums the
// squared deviations and the covariance in the same order -- the reference's
// two-pass loop, so every block's SSIM is the reference's value bit for bit
// (no contraction, IEEE double division).  The same thread adds the block's
// squared errors; pixels outside the full blocks (the right and bottom
// remainders) are spread over all threads afterwards.  Each wave reduces its
// two sums and writes them to scratch; a one-workgroup kernel then adds the
// per-wave partials in a fixed order and writes {mse, psnr, ssim}.  No
// atomics: the result is the same on every run.  The totals are sums of the
// same terms as the reference's in a different order (a tree over threads
// instead of one running sum), so they agree to the rounding of a double sum.
#include "rt_internal.h"

namespace rt {
namespace {

constexpr int kMetWave = 64;           // one wave per workgroup
constexpr uint32_t kMetMaxGroups = 8192;
constexpr int kMetFinish = 1024;       // threads of the final reduction

struct MetPlan {
    uint32_t bw, bh;   // full 8x8 blocks across / down
    uint64_t blocks;   // bw * bh
    uint64_t tail;     // pixels outside the full blocks
    uint32_t groups;   // workgroups of the partial kernel
};

MetPlan met_plan(uint32_t w, uint32_t h) {
    MetPlan p;
    p.bw = w / 8;
    p.bh = h / 8;
    p.blocks = (uint64_t)p.bw * p.bh;
    p.tail = (uint64_t)w * h - p.blocks * 64;
    const uint64_t work = p.blocks > p.tail / 64 ? p.blocks : p.tail / 64;
    uint64_t g = (work + kMetWave - 1) / kMetWave;
    g = g < 1 ? 1 : (g > kMetMaxGroups ? kMetMaxGroups : g);
    p.groups = (uint32_t)g;
    return p;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, kMetWave);
    return v;
}

template <bool VEC>
__device__ __forceinline__ void load_row(const float* __restrict__ p, float* r) {
    if (VEC) {
        const float4 x = reinterpret_cast<const float4*>(p)[0];
        const float4 y = reinterpret_cast<const float4*>(p)[1];
        r[0] = x.x; r[1] = x.y; r[2] = x.z; r[3] = x.w;
        r[4] = y.x; r[5] = y.y; r[6] = y.z; r[7] = y.w;
    } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) r[i] = p[i];
    }
}

// One 8x8 block: image_metrics.hpp:59-96 (means, n-1 variances, covariance,
// the SSIM ratio clamped to [0, 1]); sq gains the block's squared errors.
template <bool VEC>
__device__ __forceinline__ double block_ssim(const float* __restrict__ a,
                                             const float* __restrict__ b, size_t w, uint32_t bx,
                                             uint32_t by, double& sq) {
    float ra[64], rb[64];
#pragma unroll
    for (int y = 0; y < 8; ++y) {
        const size_t off = ((size_t)by * 8 + y) * w + (size_t)bx * 8;
        load_row<VEC>(a + off, ra + 8 * y);
        load_row<VEC>(b + off, rb + 8 * y);
    }
    double ma = 0.0, mb = 0.0;
#pragma unroll
    for (int i = 0; i < 64; ++i) {
        ma += ra[i];
        mb += rb[i];
        const double d = (double)ra[i] - (double)rb[i];
        sq += d * d;
    }
    ma /= 64;
    mb /= 64;
    double va = 0.0, vb = 0.0, cab = 0.0;
#pragma unroll
    for (int i = 0; i < 64; ++i) {
        const double da = ra[i] - ma;
        const double db = rb[i] - mb;
        va += da * da;
        vb += db * db;
        cab += da * db;
    }
    va /= 63;
    vb /= 63;
    cab /= 63;
    const double c1 = (0.01 * 1.0) * (0.01 * 1.0), c2 = (0.03 * 1.0) * (0.03 * 1.0);
    const double num = (2.0 * ma * mb + c1) * (2.0 * cab + c2);
    const double den = (ma * ma + mb * mb + c1) * (va + vb + c2);
    return fmin(fmax(num / den, 0.0), 1.0);
}

template <bool VEC>
__global__ __launch_bounds__(kMetWave) void k_metrics_partial(
    const float* __restrict__ a, const float* __restrict__ b, uint32_t w, uint32_t bw,
    uint32_t bh, uint64_t blocks, uint64_t tail, double* __restrict__ partial,
    double* __restrict__ block_out) {
    const uint64_t stride = (uint64_t)gridDim.x * kMetWave;
    const uint64_t gid = (uint64_t)blockIdx.x * kMetWave + threadIdx.x;
    double ss = 0.0, sq = 0.0;
    for (uint64_t k = gid; k < blocks; k += stride) {
        const uint32_t by = (uint32_t)(k / bw), bx = (uint32_t)(k % bw);
        const double s = block_ssim<VEC>(a, b, w, bx, by, sq);
        if (block_out) block_out[k] = s;
        ss += s;
    }
    // pixels right of the full blocks (rows inside them), then the rows below
    // (no full block: every pixel is in the rows below)
    const uint32_t rw = w - bw * 8, bh8 = blocks ? bh * 8 : 0;
    const uint64_t right = (uint64_t)rw * bh8;
    for (uint64_t k = gid; k < tail; k += stride) {
        uint64_t y, x;
        if (k < right) {
            y = k / rw;
            x = (uint64_t)bw * 8 + k % rw;
        } else {
            const uint64_t k2 = k - right;
            y = bh8 + k2 / w;
            x = k2 % w;
        }
        const size_t i = (size_t)y * w + x;
        const double d = (double)a[i] - (double)b[i];
        sq += d * d;
    }
    ss = wave_sum(ss);
    sq = wave_sum(sq);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = sq;
        partial[2 * blockIdx.x + 1] = ss;
    }
}

__global__ __launch_bounds__(kMetFinish) void k_metrics_finish(
    const double* __restrict__ partial, uint32_t groups, uint64_t npix, uint64_t blocks,
    double* __restrict__ out) {
    __shared__ double lsq[kMetFinish], lss[kMetFinish];
    const uint32_t t = threadIdx.x;
    double sq = 0.0, ss = 0.0;
    for (uint32_t g = t; g < groups; g += kMetFinish) {
        sq += partial[2 * g];
        ss += partial[2 * g + 1];
    }
    lsq[t] = sq;
    lss[t] = ss;
    __syncthreads();
    for (uint32_t o = kMetFinish / 2; o > 0; o >>= 1) {
        if (t < o) {
            lsq[t] += lsq[t + o];
            lss[t] += lss[t + o];
        }
        __syncthreads();
    }
    if (t == 0) {
        // image_metrics.hpp:35-40, 98-101
        const double mse = lsq[0] / (double)npix;
        out[0] = mse;
        out[1] = mse < 1e-12 ? 100.0 : 10.0 * log10(1.0 / mse);
        out[2] = blocks ? lss[0] / (double)blocks : 1.0;
    }
}

__global__ void k_metrics_empty(double* __restrict__ out) {
    if (threadIdx.x == 0) {
        out[0] = 0.0;
        out[1] = 100.0;
        out[2] = 1.0;
    }
}

}  // namespace
}  // namespace rt

using namespace rt;

extern "C" rt_status rt_image_metrics_scratch_bytes(uint32_t width, uint32_t height,
                                                    size_t* bytes) {
    if (!bytes) return RT_ERR_ARG;
    *bytes = (size_t)met_plan(width, height).groups * 2 * sizeof(double);
    return RT_OK;
}

extern "C" rt_status rt_image_metrics(rt_stream stream, const float* d_a, const float* d_b,
                                      uint32_t width, uint32_t height, double* d_out,
                                      double* d_block_ssim, void* d_scratch,
                                      size_t scratch_bytes) {
    if (!d_out) return RT_ERR_ARG;
    hipStream_t s = as_stream(stream);
    const uint64_t npix = (uint64_t)width * height;
    if (npix == 0) {
        hipLaunchKernelGGL(k_metrics_empty, dim3(1), dim3(64), 0, s, d_out);
        return launch_status("k_metrics_empty");
    }
    if (!d_a || !d_b || !d_scratch) return RT_ERR_ARG;
    const MetPlan p = met_plan(width, height);
    if (scratch_bytes < (size_t)p.groups * 2 * sizeof(double)) return RT_ERR_CAPACITY;
    if (reinterpret_cast<uintptr_t>(d_scratch) % alignof(double) ||
        reinterpret_cast<uintptr_t>(d_block_ssim) % alignof(double))
        return RT_ERR_ARG;
    // float4 rows when every block row starts on a 16-byte boundary
    const bool vec = width % 4 == 0 && reinterpret_cast<uintptr_t>(d_a) % 16 == 0 &&
                     reinterpret_cast<uintptr_t>(d_b) % 16 == 0;
    double* part = static_cast<double*>(d_scratch);
    if (vec)
        hipLaunchKernelGGL(k_metrics_partial<true>, dim3(p.groups), dim3(kMetWave), 0, s, d_a,
                           d_b, width, p.bw, p.bh, p.blocks, p.tail, part, d_block_ssim);
    else
        hipLaunchKernelGGL(k_metrics_partial<false>, dim3(p.groups), dim3(kMetWave), 0, s, d_a,
                           d_b, width, p.bw, p.bh, p.blocks, p.tail, part, d_block_ssim);
    rt_status st = launch_status("k_metrics_partial");
    if (st != RT_OK) return st;
    hipLaunchKernelGGL(k_metrics_finish, dim3(1), dim3(kMetFinish), 0, s, part, p.groups, npix,
                       p.blocks, d_out);
    return launch_status("k_metrics_finish");
}
