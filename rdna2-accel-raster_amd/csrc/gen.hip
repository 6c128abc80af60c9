// gen.hip -- synthetic inputs made on the device (meshes, ray batches) and
// the small per-element kernels either side of traversal (GpuTriangle leaf
// conversion, per-pixel light masks).  Elementwise and HBM-bound: one thread
// per output element, SoA stores coalesced across the wave.
#include <cmath>

#include "rt_internal.h"

namespace rt {
namespace {

constexpr int kB = 256;

__device__ __forceinline__ void normalize3(float& x, float& y, float& z) {
    float l = sqrtf(x * x + y * y + z * z);
    x = x / l;
    y = y / l;
    z = z / l;
}

__device__ __forceinline__ void put_ray(const rt_rays_out& o, size_t r, float ox, float oy,
                                        float oz, float tn, float dx, float dy, float dz,
                                        float tx) {
    o.ox[r] = ox; o.oy[r] = oy; o.oz[r] = oz; o.tmin[r] = tn;
    o.dx[r] = dx; o.dy[r] = dy; o.dz[r] = dz; o.tmax[r] = tx;
}

// common/mesh_gen.hpp:29-83 createTessellatedTerrain: vertex (x, y) of a res
// x res grid at (-size/2 + x step, h, -size/2 + y step) with
// h = 5 sin(.05 px) cos(.05 pz) + 2 sin(.125 px + .085 pz).  FUSED evaluates
// the three multiply-adds as the reference's -O3 -march=native build
// contracts them (fmaf); otherwise one rounding per operator.  sinf / cosf
// are the device's (within a few ulp of the host's).
template <bool FUSED>
__global__ void k_terrain_verts(uint32_t res, float step, float half, float* __restrict__ pos) {
    const uint32_t v = blockIdx.x * kB + threadIdx.x;
    if (v >= res * res) return;
    const uint32_t x = v % res, y = v / res;
    float px, pz, py;
    if (FUSED) {
        px = fmaf((float)x, step, -half);
        pz = fmaf((float)y, step, -half);
        const float fx = px * 0.05f, fz = pz * 0.05f;
        const float arg = fmaf(fx, 2.5f, fz * 1.7f);
        py = fmaf(sinf(fx) * cosf(fz), 5.0f, sinf(arg) * 2.0f);
    } else {
        px = -half + (float)x * step;
        pz = -half + (float)y * step;
        const float fx = px * 0.05f, fz = pz * 0.05f;
        py = sinf(fx) * cosf(fz) * 5.0f + sinf(fx * 2.5f + fz * 1.7f) * 2.0f;
    }
    pos[3 * (size_t)v] = px;
    pos[3 * (size_t)v + 1] = py;
    pos[3 * (size_t)v + 2] = pz;
}

// Two triangles per quad, wound (i0, i2, i1), (i1, i2, i3) (mesh_gen.hpp:70-80).
__global__ void k_terrain_idx(uint32_t res, uint32_t* __restrict__ idx) {
    const uint32_t q = blockIdx.x * kB + threadIdx.x;
    const uint32_t w = res - 1;
    if (q >= w * w) return;
    const uint32_t x = q % w, y = q / w;
    const uint32_t i0 = y * res + x, i1 = i0 + 1, i2 = i0 + res, i3 = i2 + 1;
    uint32_t* o = idx + 6 * (size_t)q;
    o[0] = i0; o[1] = i2; o[2] = i1;
    o[3] = i1; o[4] = i2; o[5] = i3;
}

// UV sphere, vertex (i, j): theta = pi j / nv, phi = 2 pi i / nu, radius
// r (1 + disp (2u - 1)); pole rows share the radius of their first vertex.
__global__ void k_sphere_verts(uint32_t nu, uint32_t nv, float radius, float disp, uint32_t seed,
                               float* __restrict__ pos) {
    const uint32_t v = blockIdx.x * kB + threadIdx.x;
    if (v >= (nv + 1) * nu) return;
    const uint32_t j = v / nu, i = v % nu;
    const float kPi = 3.14159265358979f;
    const float theta = kPi * (float)j / (float)nv;
    const float st = sinf(theta), ct = cosf(theta);
    const uint32_t hv = (j == 0 || j == nv) ? j * nu : v;
    const float phi = 2.0f * kPi * (float)i / (float)nu;
    const float r = radius * (1.0f + disp * (2.0f * hash01(seed, hv) - 1.0f));
    pos[3 * (size_t)v + 0] = r * (st * cosf(phi));
    pos[3 * (size_t)v + 1] = r * (st * sinf(phi));
    pos[3 * (size_t)v + 2] = r * ct;
}

__global__ void k_sphere_idx(uint32_t nu, uint32_t nv, uint32_t* __restrict__ idx) {
    const uint32_t q = blockIdx.x * kB + threadIdx.x;
    if (q >= nu * nv) return;
    const uint32_t j = q / nu, i = q % nu;
    const uint32_t i1 = (i + 1 == nu) ? 0 : i + 1;
    const uint32_t a = j * nu + i, b = j * nu + i1, c = (j + 1) * nu + i, d = (j + 1) * nu + i1;
    uint32_t* o = idx + 6 * (size_t)q;
    o[0] = a; o[1] = c; o[2] = b;
    o[3] = b; o[4] = c; o[5] = d;
}

// Grid topology shared by the wave grid and the heightfield: quad (c, r),
// tri 1 = v00 v10 v01, tri 2 = v10 v11 v01 (include/radv_bvh.hpp:257-279).
__global__ void k_grid_idx(uint32_t nx, uint32_t ny, uint32_t* __restrict__ idx) {
    const uint32_t q = blockIdx.x * kB + threadIdx.x;
    if (q >= nx * ny) return;
    const uint32_t r = q / nx, c = q % nx;
    const uint32_t v00 = r * (nx + 1) + c, v10 = v00 + 1, v01 = v00 + nx + 1, v11 = v01 + 1;
    uint32_t* o = idx + 6 * (size_t)q;
    o[0] = v00; o[1] = v10; o[2] = v01;
    o[3] = v10; o[4] = v11; o[5] = v01;
}

__global__ void k_wave_verts(uint32_t nx, uint32_t ny, float e, float t, float amp, float fx,
                             float fy, float* __restrict__ pos) {
    const uint32_t v = blockIdx.x * kB + threadIdx.x;
    if (v >= (nx + 1) * (ny + 1)) return;
    const uint32_t r = v / (nx + 1), c = v % (nx + 1);
    const float dx = (2.0f * e) / (float)nx, dy = (2.0f * e) / (float)ny;
    const float x = -e + (float)c * dx, y = -e + (float)r * dy;
    pos[3 * (size_t)v + 0] = x;
    pos[3 * (size_t)v + 1] = y;
    pos[3 * (size_t)v + 2] = amp * sinf(fx * x + t) * cosf(fy * y);
}

__global__ void k_height_verts(uint32_t nx, uint32_t ny, float e, float amp, uint32_t seed,
                               float* __restrict__ pos) {
    const uint32_t v = blockIdx.x * kB + threadIdx.x;
    if (v >= (nx + 1) * (ny + 1)) return;
    const uint32_t r = v / (nx + 1), c = v % (nx + 1);
    const float dx = (2.0f * e) / (float)nx, dy = (2.0f * e) / (float)ny;
    pos[3 * (size_t)v + 0] = -e + (float)c * dx;
    pos[3 * (size_t)v + 1] = -e + (float)r * dy;
    pos[3 * (size_t)v + 2] = amp * (2.0f * hash01(seed, v) - 1.0f);
}

// include/radv_bvh.hpp:224-283 generate_dynamic_grid_mesh, one quad per
// thread (cols / rows chosen on the host exactly as the reference does).
__global__ void k_dynamic_grid(uint32_t cols, uint32_t rows, float t, float extent,
                               float* __restrict__ xyz) {
    const uint32_t q = blockIdx.x * kB + threadIdx.x;
    if (q >= cols * rows) return;
    dyn_grid_quad(cols, rows, t, extent, q, xyz + 18 * (size_t)q);
}

// Many deforming grids in one launch: grid g's quads are blockIdx.y = g.
__global__ void k_dynamic_grid_batch(const GridGen* __restrict__ gen) {
    const GridGen g = gen[blockIdx.y];
    const uint32_t q = blockIdx.x * kB + threadIdx.x;
    if (q >= g.cols * g.rows) return;
    dyn_grid_quad(g.cols, g.rows, g.t, g.extent, q, g.xyz + 18 * (size_t)q);
}

struct CamBasis {
    float eye[3], F[3], R[3], U[3];
    float sxs, sys, tmin, tmax;
};

__global__ void k_camera(uint32_t w, uint32_t h, CamBasis cb, rt_rays_out o) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= w * h) return;
    const uint32_t x = i % w, y = i / w;
    const float sx = ((float)x + 0.5f) * cb.sxs - 1.0f;
    const float sy = 1.0f - ((float)y + 0.5f) * cb.sys;
    float d0 = cb.F[0] + sx * cb.R[0] + sy * cb.U[0];
    float d1 = cb.F[1] + sx * cb.R[1] + sy * cb.U[1];
    float d2 = cb.F[2] + sx * cb.R[2] + sy * cb.U[2];
    normalize3(d0, d1, d2);
    put_ray(o, i, cb.eye[0], cb.eye[1], cb.eye[2], cb.tmin, d0, d1, d2, cb.tmax);
}

// Unit normal of triangle prim (indexed, or 9 floats per triangle when idx
// is null), turned to face against the direction d.
__device__ __forceinline__ void face_normal(const float* __restrict__ pos,
                                            const uint32_t* __restrict__ idx, uint32_t prim,
                                            float ddx, float ddy, float ddz, float n[3]) {
    const float *p0, *p1, *p2;
    if (idx) {
        p0 = pos + 3 * (size_t)idx[3 * (size_t)prim + 0];
        p1 = pos + 3 * (size_t)idx[3 * (size_t)prim + 1];
        p2 = pos + 3 * (size_t)idx[3 * (size_t)prim + 2];
    } else {
        p0 = pos + 9 * (size_t)prim;
        p1 = p0 + 3;
        p2 = p0 + 6;
    }
    const float e1x = p1[0] - p0[0], e1y = p1[1] - p0[1], e1z = p1[2] - p0[2];
    const float e2x = p2[0] - p0[0], e2y = p2[1] - p0[1], e2z = p2[2] - p0[2];
    float nx = e1y * e2z - e1z * e2y;
    float ny = e1z * e2x - e1x * e2z;
    float nz = e1x * e2y - e1y * e2x;
    normalize3(nx, ny, nz);
    if (nx * ddx + ny * ddy + nz * ddz > 0.0f) {
        nx = -nx;
        ny = -ny;
        nz = -nz;
    }
    n[0] = nx; n[1] = ny; n[2] = nz;
}

// Shadow rays (role of tiled_lighting_ra.comp:136-155): origin P + eps N,
// towards each light, [eps, distance).  Light-major: ray l * npix + p.
__global__ void k_shadow(const float* __restrict__ pos, const uint32_t* __restrict__ idx,
                         rt_rays_soa pr, const uint32_t* __restrict__ prim,
                         const float* __restrict__ tt, uint32_t npix,
                         const float* __restrict__ lights, uint32_t nl, float eps,
                         rt_rays_out o) {
    const uint32_t p = blockIdx.x * kB + threadIdx.x;
    if (p >= npix) return;
    const uint32_t hp = prim[p];
    if (hp == RT_MISS) {
        for (uint32_t l = 0; l < nl; ++l)
            put_ray(o, (size_t)l * npix + p, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 1.f, 0.f);
        return;
    }
    const float dx = pr.dx[p], dy = pr.dy[p], dz = pr.dz[p];
    float n[3];
    face_normal(pos, idx, hp, dx, dy, dz, n);
    const float t = tt[p];
    const float Px = pr.ox[p] + t * dx, Py = pr.oy[p] + t * dy, Pz = pr.oz[p] + t * dz;
    const float Ox = Px + eps * n[0], Oy = Py + eps * n[1], Oz = Pz + eps * n[2];
    for (uint32_t l = 0; l < nl; ++l) {
        const float Lx = lights[3 * l + 0] - Ox;
        const float Ly = lights[3 * l + 1] - Oy;
        const float Lz = lights[3 * l + 2] - Oz;
        const float dist = sqrtf(Lx * Lx + Ly * Ly + Lz * Lz);
        put_ray(o, (size_t)l * npix + p, Ox, Oy, Oz, eps, Lx / dist, Ly / dist, Lz / dist, dist);
    }
}

// shaders/rq_ao.comp:18-25 hash
__device__ __forceinline__ uint32_t ao_hash(uint32_t x) {
    x += (x << 10u); x ^= (x >> 6u); x += (x << 3u); x ^= (x >> 11u); x += (x << 15u);
    return x;
}

// One cosine-weighted hemisphere ray per pixel (rq_ao.comp:32-45,70-87).
// One cosine-weighted hemisphere ray (rq_ao.comp:32-97) about the unit
// normal n for pixel p of a w x h frame, from P + offset n.
__device__ __forceinline__ void ao_ray(uint32_t p, const float n[3], float Px, float Py,
                                       float Pz, uint32_t w, uint32_t h, uint32_t frame,
                                       float offset, float tmin, float tmax, rt_rays_out o) {
    uint32_t seed = p + 5544u + frame * w * h;
    seed = ao_hash(seed);
    const float u1 = (float)seed / 4294967295.0f;
    seed = ao_hash(seed);
    const float u2 = (float)seed / 4294967295.0f;
    const float r = sqrtf(u1);
    const float th = 6.2831853f * u2;
    const float lx = r * cosf(th), ly = r * sinf(th);
    const float lz = sqrtf(fmaxf(0.0f, 1.0f - u1));
    float upx = 0.0f, upy = 0.0f, upz = 1.0f;
    if (!(fabsf(n[2]) < 0.999f)) {
        upx = 1.0f;
        upz = 0.0f;
    }
    float tx0 = upy * n[2] - upz * n[1];
    float ty0 = upz * n[0] - upx * n[2];
    float tz0 = upx * n[1] - upy * n[0];
    normalize3(tx0, ty0, tz0);
    const float bx = n[1] * tz0 - n[2] * ty0;
    const float by = n[2] * tx0 - n[0] * tz0;
    const float bz = n[0] * ty0 - n[1] * tx0;
    float d0 = tx0 * lx + bx * ly + n[0] * lz;
    float d1 = ty0 * lx + by * ly + n[1] * lz;
    float d2 = tz0 * lx + bz * ly + n[2] * lz;
    normalize3(d0, d1, d2);
    put_ray(o, p, Px + offset * n[0], Py + offset * n[1], Pz + offset * n[2], tmin, d0, d1, d2,
            tmax);
}

__global__ void k_ao(const float* __restrict__ pos, const uint32_t* __restrict__ idx,
                     rt_rays_soa pr, const uint32_t* __restrict__ prim,
                     const float* __restrict__ tt, uint32_t w, uint32_t h, uint32_t frame,
                     float offset, float tmin, float tmax, rt_rays_out o) {
    const uint32_t p = blockIdx.x * kB + threadIdx.x;
    if (p >= w * h) return;
    const uint32_t hp = prim[p];
    if (hp == RT_MISS) {
        put_ray(o, p, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 1.f, 0.f);
        return;
    }
    const float dx = pr.dx[p], dy = pr.dy[p], dz = pr.dz[p];
    float n[3];
    face_normal(pos, idx, hp, dx, dy, dz, n);
    const float t = tt[p];
    ao_ray(p, n, pr.ox[p] + t * dx, pr.oy[p] + t * dy, pr.oz[p] + t * dz, w, h, frame, offset,
           tmin, tmax, o);
}

// AO rays over a two-level scene: the normal of the hit triangle of
// instance inst[p], oriented against the mapped direction in object space,
// then mapped to world space by M^T (the inverse transpose of the
// object-to-world map) for a general instance.
__global__ void k_ao_scene(const rt_instance* __restrict__ inst, rt_rays_soa pr,
                           const uint32_t* __restrict__ prim, const uint32_t* __restrict__ hit_inst,
                           const float* __restrict__ tt, uint32_t w, uint32_t h, uint32_t frame,
                           float offset, float tmin, float tmax, rt_rays_out o) {
    const uint32_t p = blockIdx.x * kB + threadIdx.x;
    if (p >= w * h) return;
    const uint32_t hp = prim[p];
    if (hp == RT_MISS) {
        put_ray(o, p, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 1.f, 0.f);
        return;
    }
    const rt_instance& I = inst[hit_inst[p]];
    const float dx = pr.dx[p], dy = pr.dy[p], dz = pr.dz[p];
    float n[3];
    if (I.kind == RT_INST_GENERAL) {
        const float* m = I.world_to_object;
        const float ex = m[0] * dx + m[1] * dy + m[2] * dz;
        const float ey = m[4] * dx + m[5] * dy + m[6] * dz;
        const float ez = m[8] * dx + m[9] * dy + m[10] * dz;
        float q[3];
        face_normal(I.pos, I.idx, hp, ex, ey, ez, q);
        n[0] = m[0] * q[0] + m[4] * q[1] + m[8] * q[2];
        n[1] = m[1] * q[0] + m[5] * q[1] + m[9] * q[2];
        n[2] = m[2] * q[0] + m[6] * q[1] + m[10] * q[2];
        normalize3(n[0], n[1], n[2]);
    } else {
        face_normal(I.pos, I.idx, hp, dx, dy, dz, n);
    }
    const float t = tt[p];
    ao_ray(p, n, pr.ox[p] + t * dx, pr.oy[p] + t * dy, pr.oz[p] + t * dz, w, h, frame, offset,
           tmin, tmax, o);
}

// rq_ao.comp's 1 SPP resolve: 1 for a primary miss, else 0 / 1 by occlusion.
__global__ void k_ao_resolve(const uint32_t* __restrict__ bits, const uint32_t* __restrict__ prim,
                             uint32_t npix, float* __restrict__ ao) {
    const uint32_t p = blockIdx.x * kB + threadIdx.x;
    if (p >= npix) return;
    const bool occ = (bits[p >> 5] >> (p & 31u)) & 1u;
    ao[p] = prim[p] == RT_MISS ? 1.0f : (occ ? 0.0f : 1.0f);
}

// 7x7 depth-gated bilateral filter over a 16 x 16 tile per workgroup, the
// tile and its 3-pixel apron staged in LDS.  Taps accumulate row by row,
// left to right, as the CPU filter does, so sums round identically.
constexpr int kBT = 16;
__global__ __launch_bounds__(kBT * kBT) void k_blur7(const float* __restrict__ in,
                                                      const float* __restrict__ depth,
                                                      uint32_t w, uint32_t h, float thr,
                                                      float* __restrict__ out) {
    constexpr int A = kBT + 6;
    __shared__ float si[A][A], sd[A][A];
    const int x0 = (int)(blockIdx.x * kBT) - 3, y0 = (int)(blockIdx.y * kBT) - 3;
    for (int k = threadIdx.y * kBT + threadIdx.x; k < A * A; k += kBT * kBT) {
        const int ly = k / A, lx = k % A, gx = x0 + lx, gy = y0 + ly;
        const bool inside = gx >= 0 && gy >= 0 && gx < (int)w && gy < (int)h;
        si[ly][lx] = inside ? in[(size_t)gy * w + gx] : 0.0f;
        sd[ly][lx] = inside ? depth[(size_t)gy * w + gx] : 0.0f;
    }
    __syncthreads();
    const int x = (int)(blockIdx.x * kBT + threadIdx.x), y = (int)(blockIdx.y * kBT + threadIdx.y);
    if (x >= (int)w || y >= (int)h) return;
    const float cd = sd[threadIdx.y + 3][threadIdx.x + 3];
    if (cd >= 1e30f) {
        out[(size_t)y * w + x] = si[threadIdx.y + 3][threadIdx.x + 3];
        return;
    }
    const float wt[4] = {0.227027f, 0.1945946f, 0.1216216f, 0.054054f};
    float acc = 0.0f, tot = 0.0f;
    for (int j = -3; j <= 3; ++j) {
        if (y + j < 0 || y + j >= (int)h) continue;
        for (int i = -3; i <= 3; ++i) {
            if (x + i < 0 || x + i >= (int)w) continue;
            const float dq = sd[threadIdx.y + 3 + j][threadIdx.x + 3 + i];
            if (!(fabsf(dq - cd) < thr)) continue;
            const float ww = wt[i < 0 ? -i : i] * wt[j < 0 ? -j : j];
            acc = acc + si[threadIdx.y + 3 + j][threadIdx.x + 3 + i] * ww;
            tot = tot + ww;
        }
    }
    out[(size_t)y * w + x] = acc / fmaxf(tot, 0.001f);
}

__global__ void k_random_rays(uint32_t count, uint32_t seed, float lx, float ly, float lz,
                              float hx, float hy, float hz, float tmax, rt_rays_out o) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= count) return;
    const float lo[3] = {lx, ly, lz}, hi[3] = {hx, hy, hz};
    float org[3], dir[3];
    random_ray(seed, i, lo, hi, org, dir);
    put_ray(o, i, org[0], org[1], org[2], 0.0f, dir[0], dir[1], dir[2], tmax);
}

// The synthetic payloads of many ray requests in one launch: over the chunk
// table of a batched trace (item i owns chunks [chunk[i], chunk[i + 1]) of
// chunk_rays rays), ray r of item i is ray r of rt_gen_random_rays with that
// item's generator (seed, box, tmax), written to the item's rays.
__global__ void k_random_rays_batch(const rt_trace_item* __restrict__ items,
                                    const rt_ray_gen* __restrict__ gen,
                                    const uint32_t* __restrict__ chunk, uint32_t count,
                                    uint32_t total_chunks, uint32_t chunk_rays) {
    const uint64_t g = (uint64_t)blockIdx.x * kB + threadIdx.x;
    const uint32_t c = (uint32_t)(g / chunk_rays);
    if (c >= total_chunks) return;
    uint32_t lo = 0, hi = count;  // last item with chunk[item] <= c
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (chunk[mid] <= c) lo = mid;
        else hi = mid;
    }
    const rt_trace_item& it = items[lo];
    const uint32_t r = (c - chunk[lo]) * chunk_rays + (uint32_t)(g % chunk_rays);
    if (r >= it.count) return;
    const rt_ray_gen& q = gen[lo];
    float org[3], dir[3];
    random_ray(q.seed, r, q.lo, q.hi, org, dir);
    const rt_rays_out o = {const_cast<float*>(it.rays.ox), const_cast<float*>(it.rays.oy),
                           const_cast<float*>(it.rays.oz), const_cast<float*>(it.rays.tmin),
                           const_cast<float*>(it.rays.dx), const_cast<float*>(it.rays.dy),
                           const_cast<float*>(it.rays.dz), const_cast<float*>(it.rays.tmax)};
    put_ray(o, r, org[0], org[1], org[2], 0.0f, dir[0], dir[1], dir[2], q.tmax);
}

// software_bvh_bvh4.comp:50-94 procedural rays of the exp0 probe.
// Coherent: ray i is pixel (i % 1024, i / 1024) of a 1024-wide grid seen from
// cam towards look, dir = normalize(fwd + right u 0.55 + up v 0.55) with the
// basis fwd = normalize(look - cam), right = normalize(cross(fwd, +y)), up =
// cross(right, fwd).  Incoherent: origin and direction from six pcg_hash
// values of i (hashFloat = float(pcg_hash(6 i + k)) * float(1 / (2^32 - 1))):
// origin in [-40,40] x [15,30] x [-40,40], dir = normalize(U(-.4,.4),
// -1 - U(0,.5), U(-.4,.4)).  Products are rounded one operator at a time in
// the shader's left-to-right order (the basis is the same for every ray and
// passed in, computed on the host the same way).
struct Exp0Basis {
    float F[3], R[3], U[3], cam[3];
    float tmin, tmax;
    uint32_t height;
};

__global__ void k_exp0_rays(uint32_t count, int coherent, Exp0Basis b, rt_rays_out o) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= count) return;
    float o0, o1, o2, d0, d1, d2;
    if (coherent) {
        const uint32_t x = i % 1024u, y = i / 1024u;
        const float u = ((float)x + 0.5f) / 1024.0f * 2.0f - 1.0f;
        const float v = ((float)y + 0.5f) / (float)b.height * 2.0f - 1.0f;
        o0 = b.cam[0]; o1 = b.cam[1]; o2 = b.cam[2];
        d0 = b.F[0] + b.R[0] * u * 0.55f + b.U[0] * v * 0.55f;
        d1 = b.F[1] + b.R[1] * u * 0.55f + b.U[1] * v * 0.55f;
        d2 = b.F[2] + b.R[2] * u * 0.55f + b.U[2] * v * 0.55f;
    } else {
        const float hs = 1.0f / 4294967295.0f;
        const uint32_t s = i * 6u;
        o0 = (float)pcg_hash(s + 0u) * hs * 80.0f - 40.0f;
        o1 = (float)pcg_hash(s + 1u) * hs * 15.0f + 15.0f;
        o2 = (float)pcg_hash(s + 2u) * hs * 80.0f - 40.0f;
        d0 = (float)pcg_hash(s + 3u) * hs * 0.8f - 0.4f;
        d1 = -1.0f - (float)pcg_hash(s + 4u) * hs * 0.5f;
        d2 = (float)pcg_hash(s + 5u) * hs * 0.8f - 0.4f;
    }
    normalize3(d0, d1, d2);
    put_ray(o, i, o0, o1, o2, b.tmin, d0, d1, d2, b.tmax);
}

__global__ void k_convert48(const float* __restrict__ t48, const uint32_t* __restrict__ order,
                            uint32_t n, rt_tri64* __restrict__ out) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const float4* s = reinterpret_cast<const float4*>(t48) + 3 * (size_t)i;
    const float4 a = s[0], b = s[1], c = s[2];
    float4* d = reinterpret_cast<float4*>(out + i);
    d[0] = make_float4(a.x, a.y, a.z, b.x);
    d[1] = make_float4(b.y, b.z, c.x, c.y);
    d[2] = make_float4(c.z, 0.f, 0.f, 0.f);
    reinterpret_cast<uint4*>(d)[3] = make_uint4(order ? order[i] : i, 0x80000000u, 0u, 9u);
}

__global__ void k_pack_masks(const uint32_t* __restrict__ bits, uint32_t npix, uint32_t nl,
                             uint16_t* __restrict__ masks) {
    const uint32_t p = blockIdx.x * kB + threadIdx.x;
    if (p >= npix) return;
    uint32_t m = 0;
    for (uint32_t l = 0; l < nl; ++l) {
        const uint64_t r = (uint64_t)l * npix + p;
        m |= ((bits[r >> 5] >> (r & 31u)) & 1u) << l;
    }
    masks[p] = (uint16_t)m;
}

}  // namespace
}  // namespace rt

using namespace rt;

extern "C" rt_status rt_gen_sphere(rt_stream stream, uint32_t nu, uint32_t nv, float radius,
                                   float disp, uint32_t seed, float* d_pos, uint32_t* d_idx) {
    if (!d_pos || !d_idx || nu < 3 || nv < 2) return RT_ERR_ARG;
    if ((uint64_t)nu * nv * 2 > 0xFFFFFFFFull) return RT_ERR_CAPACITY;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_sphere_verts, dim3(ceil_div((uint64_t)(nv + 1) * nu, kB)), dim3(kB), 0, s,
                       nu, nv, radius, disp, seed, d_pos);
    hipLaunchKernelGGL(k_sphere_idx, dim3(ceil_div((uint64_t)nu * nv, kB)), dim3(kB), 0, s, nu, nv,
                       d_idx);
    return launch_status("rt_gen_sphere");
}

extern "C" rt_status rt_gen_wave_grid(rt_stream stream, uint32_t nx, uint32_t ny, float extent,
                                      float t, float amp, float fx, float fy, float* d_pos,
                                      uint32_t* d_idx) {
    if (!d_pos || !d_idx || nx == 0 || ny == 0) return RT_ERR_ARG;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_wave_verts, dim3(ceil_div((uint64_t)(nx + 1) * (ny + 1), kB)), dim3(kB),
                       0, s, nx, ny, extent, t, amp, fx, fy, d_pos);
    hipLaunchKernelGGL(k_grid_idx, dim3(ceil_div((uint64_t)nx * ny, kB)), dim3(kB), 0, s, nx, ny,
                       d_idx);
    return launch_status("rt_gen_wave_grid");
}

extern "C" rt_status rt_gen_heightfield(rt_stream stream, uint32_t nx, uint32_t ny, float extent,
                                        float amp, uint32_t seed, float* d_pos, uint32_t* d_idx) {
    if (!d_pos || !d_idx || nx == 0 || ny == 0) return RT_ERR_ARG;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_height_verts, dim3(ceil_div((uint64_t)(nx + 1) * (ny + 1), kB)),
                       dim3(kB), 0, s, nx, ny, extent, amp, seed, d_pos);
    hipLaunchKernelGGL(k_grid_idx, dim3(ceil_div((uint64_t)nx * ny, kB)), dim3(kB), 0, s, nx, ny,
                       d_idx);
    return launch_status("rt_gen_heightfield");
}

extern "C" rt_status rt_gen_dynamic_grid(rt_stream stream, uint32_t n, float t, float extent,
                                         float* d_xyz9) {
    if (!d_xyz9 || n < 2) return RT_ERR_ARG;
    uint32_t cols, rows;
    dyn_grid_shape(n, &cols, &rows);
    hipLaunchKernelGGL(k_dynamic_grid, dim3(ceil_div(cols * rows, kB)), dim3(kB), 0,
                       as_stream(stream), cols, rows, t, extent, d_xyz9);
    return launch_status("rt_gen_dynamic_grid");
}

namespace rt {
rt_status launch_grid_gen(hipStream_t s, const GridGen* d_gen, uint32_t count,
                          uint32_t max_quads) {
    if (count == 0 || max_quads == 0) return RT_OK;
    hipLaunchKernelGGL(k_dynamic_grid_batch, dim3(ceil_div(max_quads, kB), count), dim3(kB), 0, s,
                       d_gen);
    return launch_status("k_dynamic_grid_batch");
}

rt_status launch_ray_gen(hipStream_t s, const rt_trace_item* d_items, const rt_ray_gen* d_gen,
                         const uint32_t* d_chunk_start, uint32_t count, uint32_t total_chunks,
                         uint32_t chunk_rays) {
    if (count == 0 || total_chunks == 0) return RT_OK;
    hipLaunchKernelGGL(k_random_rays_batch,
                       dim3(ceil_div((uint64_t)total_chunks * chunk_rays, kB)), dim3(kB), 0, s,
                       d_items, d_gen, d_chunk_start, count, total_chunks, chunk_rays);
    return launch_status("k_random_rays_batch");
}
}  // namespace rt

// common/mesh_gen.hpp:31-34: res = floor(sqrt(target / 2)) + 1, at least 2.
static uint32_t terrain_res(uint32_t target) {
    const uint32_t res = (uint32_t)std::sqrt((double)(target / 2)) + 1u;
    return res < 2u ? 2u : res;
}

extern "C" rt_status rt_terrain_resolution(uint32_t target, uint32_t* res, uint32_t* nverts,
                                           uint32_t* ntris) {
    const uint32_t r = terrain_res(target);
    if ((uint64_t)r * r > 0xFFFFFFFFull) return RT_ERR_CAPACITY;
    if (res) *res = r;
    if (nverts) *nverts = r * r;
    if (ntris) *ntris = (r - 1) * (r - 1) * 2;
    return RT_OK;
}

extern "C" rt_status rt_gen_terrain(rt_stream stream, uint32_t target, float size, int fused,
                                    float* d_pos, uint32_t* d_idx) {
    if (!d_pos || !d_idx) return RT_ERR_ARG;
    const uint32_t res = terrain_res(target);
    if ((uint64_t)res * res > 0xFFFFFFFFull / 3) return RT_ERR_CAPACITY;
    const float step = size / (float)(res - 1), half = size * 0.5f;
    hipStream_t s = as_stream(stream);
    const uint32_t nv = res * res, nq = (res - 1) * (res - 1);
    if (fused)
        hipLaunchKernelGGL(k_terrain_verts<true>, dim3(ceil_div(nv, kB)), dim3(kB), 0, s, res,
                           step, half, d_pos);
    else
        hipLaunchKernelGGL(k_terrain_verts<false>, dim3(ceil_div(nv, kB)), dim3(kB), 0, s, res,
                           step, half, d_pos);
    hipLaunchKernelGGL(k_terrain_idx, dim3(ceil_div(nq, kB)), dim3(kB), 0, s, res, d_idx);
    return launch_status("rt_gen_terrain");
}

extern "C" rt_status rt_gen_camera_rays(rt_stream stream, uint32_t w, uint32_t h,
                                        const float eye[3], const float look[3],
                                        const float up[3], float vfov_deg, float tmin,
                                        float tmax, const rt_rays_out* out) {
    if (!eye || !look || !up || !out || w == 0 || h == 0) return RT_ERR_ARG;
    // Basis in double, rounded to float once.
    double f[3] = {(double)look[0] - eye[0], (double)look[1] - eye[1], (double)look[2] - eye[2]};
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& c : f) c /= fl;
    double r[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2],
                   f[0] * up[1] - f[1] * up[0]};
    double rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    if (!(fl > 0.0) || !(rl > 0.0)) return RT_ERR_ARG;
    for (double& c : r) c /= rl;
    double u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2],
                   r[0] * f[1] - r[1] * f[0]};
    double th = std::tan(vfov_deg * 3.14159265358979323846 / 360.0);
    double aspect = (double)w / (double)h;
    CamBasis cb;
    for (int c = 0; c < 3; ++c) {
        cb.eye[c] = eye[c];
        cb.F[c] = (float)f[c];
        cb.R[c] = (float)(r[c] * th * aspect);
        cb.U[c] = (float)(u[c] * th);
    }
    cb.sxs = 2.0f / (float)w;
    cb.sys = 2.0f / (float)h;
    cb.tmin = tmin;
    cb.tmax = tmax;
    hipLaunchKernelGGL(k_camera, dim3(ceil_div((uint64_t)w * h, kB)), dim3(kB), 0,
                       as_stream(stream), w, h, cb, *out);
    return launch_status("rt_gen_camera_rays");
}

extern "C" rt_status rt_gen_shadow_rays(rt_stream stream, const float* d_pos,
                                        const uint32_t* d_idx, const rt_rays_soa* primary,
                                        const uint32_t* d_prim, const float* d_t, uint32_t npix,
                                        const float* d_lights, uint32_t nlights, float eps,
                                        const rt_rays_out* out) {
    if (!d_pos || !d_idx || !primary || !d_prim || !d_t || !d_lights || !out || nlights == 0)
        return RT_ERR_ARG;
    if (npix == 0) return RT_OK;
    if ((uint64_t)npix * nlights > 0xFFFFFFFFull) return RT_ERR_CAPACITY;
    hipLaunchKernelGGL(k_shadow, dim3(ceil_div(npix, kB)), dim3(kB), 0, as_stream(stream), d_pos,
                       d_idx, *primary, d_prim, d_t, npix, d_lights, nlights, eps, *out);
    return launch_status("rt_gen_shadow_rays");
}

extern "C" rt_status rt_gen_ao_rays(rt_stream stream, const float* d_pos, const uint32_t* d_idx,
                                    const rt_rays_soa* primary, const uint32_t* d_prim,
                                    const float* d_t, uint32_t w, uint32_t h, uint32_t frame,
                                    float offset, float tmin, float tmax,
                                    const rt_rays_out* out) {
    if (!d_pos || !primary || !d_prim || !d_t || !out) return RT_ERR_ARG;  // d_idx may be null
    if ((uint64_t)w * h == 0) return RT_OK;
    hipLaunchKernelGGL(k_ao, dim3(ceil_div((uint64_t)w * h, kB)), dim3(kB), 0, as_stream(stream),
                       d_pos, d_idx, *primary, d_prim, d_t, w, h, frame, offset, tmin, tmax, *out);
    return launch_status("rt_gen_ao_rays");
}

extern "C" rt_status rt_gen_ao_rays_scene(rt_stream stream, const rt_instance* d_instances,
                                          uint32_t count, const rt_rays_soa* primary,
                                          const uint32_t* d_prim, const uint32_t* d_inst,
                                          const float* d_t, uint32_t w, uint32_t h,
                                          uint32_t frame, float offset, float tmin, float tmax,
                                          const rt_rays_out* out) {
    if (!d_instances || count == 0 || !primary || !d_prim || !d_inst || !d_t || !out)
        return RT_ERR_ARG;
    if ((uint64_t)w * h == 0) return RT_OK;
    hipLaunchKernelGGL(k_ao_scene, dim3(ceil_div((uint64_t)w * h, kB)), dim3(kB), 0,
                       as_stream(stream), d_instances, *primary, d_prim, d_inst, d_t, w, h, frame,
                       offset, tmin, tmax, *out);
    return launch_status("rt_gen_ao_rays_scene");
}

extern "C" rt_status rt_ao_resolve(rt_stream stream, const uint32_t* d_bits,
                                   const uint32_t* d_prim, uint32_t npix, float* d_ao) {
    if (npix == 0) return RT_OK;
    if (!d_bits || !d_prim || !d_ao) return RT_ERR_ARG;
    hipLaunchKernelGGL(k_ao_resolve, dim3(ceil_div(npix, kB)), dim3(kB), 0, as_stream(stream),
                       d_bits, d_prim, npix, d_ao);
    return launch_status("rt_ao_resolve");
}

extern "C" rt_status rt_blur7(rt_stream stream, const float* d_in, const float* d_depth,
                              uint32_t w, uint32_t h, float depth_thr, float* d_out) {
    if ((uint64_t)w * h == 0) return RT_OK;
    if (!d_in || !d_depth || !d_out || d_in == d_out) return RT_ERR_ARG;
    hipLaunchKernelGGL(k_blur7, dim3(ceil_div(w, kBT), ceil_div(h, kBT)), dim3(kBT, kBT), 0,
                       as_stream(stream), d_in, d_depth, w, h, depth_thr, d_out);
    return launch_status("rt_blur7");
}

extern "C" rt_status rt_gen_random_rays(rt_stream stream, uint32_t count, uint32_t seed,
                                        const float lo[3], const float hi[3], float tmax,
                                        const rt_rays_out* out) {
    if (!lo || !hi || !out) return RT_ERR_ARG;
    if (count == 0) return RT_OK;
    hipLaunchKernelGGL(k_random_rays, dim3(ceil_div(count, kB)), dim3(kB), 0, as_stream(stream),
                       count, seed, lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], tmax, *out);
    return launch_status("rt_gen_random_rays");
}

extern "C" rt_status rt_gen_exp0_rays(rt_stream stream, uint32_t count, int coherent,
                                      const float cam[3], const float look[3], float tmin,
                                      float tmax, const rt_rays_out* out) {
    if (!cam || !look || !out) return RT_ERR_ARG;
    if (count == 0) return RT_OK;
    // the shader's basis (software_bvh_bvh4.comp:76-78), float operators in order
    Exp0Basis b;
    float fx = look[0] - cam[0], fy = look[1] - cam[1], fz = look[2] - cam[2];
    float l = std::sqrt(fx * fx + fy * fy + fz * fz);
    fx = fx / l; fy = fy / l; fz = fz / l;
    // cross(fwd, (0, 1, 0))
    float rx = fy * 0.0f - fz * 1.0f, ry = fz * 0.0f - fx * 0.0f, rz = fx * 1.0f - fy * 0.0f;
    l = std::sqrt(rx * rx + ry * ry + rz * rz);
    rx = rx / l; ry = ry / l; rz = rz / l;
    b.F[0] = fx; b.F[1] = fy; b.F[2] = fz;
    b.R[0] = rx; b.R[1] = ry; b.R[2] = rz;
    b.U[0] = ry * fz - rz * fy; b.U[1] = rz * fx - rx * fz; b.U[2] = rx * fy - ry * fx;
    for (int k = 0; k < 3; ++k) b.cam[k] = cam[k];
    b.tmin = tmin;
    b.tmax = tmax;
    b.height = (uint32_t)(((uint64_t)count + 1023u) / 1024u);
    hipLaunchKernelGGL(k_exp0_rays, dim3(ceil_div(count, kB)), dim3(kB), 0, as_stream(stream),
                       count, coherent, b, *out);
    return launch_status("rt_gen_exp0_rays");
}

extern "C" rt_status rt_convert_tris48(rt_stream stream, const float* d_tris48,
                                       const uint32_t* d_order, uint32_t n, rt_tri64* d_out) {
    if (n == 0) return RT_OK;
    if (!d_tris48 || !d_out) return RT_ERR_ARG;
    hipLaunchKernelGGL(k_convert48, dim3(ceil_div(n, kB)), dim3(kB), 0, as_stream(stream),
                       d_tris48, d_order, n, d_out);
    return launch_status("rt_convert_tris48");
}

extern "C" rt_status rt_pack_light_masks(rt_stream stream, const uint32_t* d_bits, uint32_t npix,
                                         uint32_t nlights, uint16_t* d_masks) {
    if (!d_bits || !d_masks || nlights == 0 || nlights > 16) return RT_ERR_ARG;
    if (npix == 0) return RT_OK;
    hipLaunchKernelGGL(k_pack_masks, dim3(ceil_div(npix, kB)), dim3(kB), 0, as_stream(stream),
                       d_bits, npix, nlights, d_masks);
    return launch_status("rt_pack_light_masks");
}
