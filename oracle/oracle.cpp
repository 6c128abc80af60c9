// oracle.cpp -- CPU restatement of the reference BVH build-and-traverse path.
//
// TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
// product library.  Compiled with -ffp-contract=off: every float expression
// rounds once per operator, in the order written.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

inline float fmin3(float a, float b, float c) { return std::min(a, std::min(b, c)); }
inline float fmax3(float a, float b, float c) { return std::max(a, std::max(b, c)); }

// Uniform in [0,1) with 24 random bits: exact on any IEEE float unit.
inline float hash01(uint32_t seed, uint32_t i) {
    uint32_t h = or_pcg_hash(i * 0x9E3779B9u + seed);
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}

// least e in [lo, hi] with pred(e) true, pred monotone false..true; hi if none
template <class P>
static int least_true(int lo, int hi, P pred) {
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (pred(mid)) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

template <class F>
void parallel_chunks(uint32_t n, int nthreads, F fn) {
    const uint32_t chunk = 4096;  // rays handed to a thread at a time
    if (nthreads <= 1 || n <= chunk) {
        fn(0u, n);
        return;
    }
    std::vector<std::thread> pool;
    std::atomic<uint32_t> next{0};
    for (int t = 0; t < nthreads; ++t) {
        pool.emplace_back([&]() {
            for (;;) {
                uint32_t b = next.fetch_add(chunk);
                if (b >= n) break;
                fn(b, std::min(n, b + chunk));
            }
        });
    }
    for (auto& th : pool) th.join();
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// Mesh generators
// ---------------------------------------------------------------------------

// common/mesh_gen.hpp:31-34 (grid resolution from a triangle target)
uint32_t or_terrain_res(uint32_t target) {
    uint32_t quads = target / 2;
    uint32_t res = (uint32_t)std::sqrt((double)quads) + 1;  // std::sqrt(uint) -> double
    if (res < 2) res = 2;
    return res;
}

// common/mesh_gen.hpp:29-83 createTessellatedTerrain: deterministic heightfield,
// y = 5 sin(.05x) cos(.05z) + 2 sin(.125x + .085z), two triangles per quad
// wound (i0,i2,i1), (i1,i2,i3).
//
// fused = 0 evaluates the source expressions with one rounding per operator.
// fused = 1 evaluates them as the reference's own build does: CMakeLists.txt:9
// compiles with -O3 -march=native, under which g++ contracts
//   px = fma(x, step, -half),  arg = fma(fx, 2.5, fz * 1.7),
//   py = fma(sin(fx) cos(fz), 5, sin(arg) * 2)
// (established bit-exact against the headers built with the reference's
// flags).  The published exp0_ray_stats.csv counters come from that build.
void or_terrain_ex(uint32_t target, float size, int fused, float* pos, uint32_t* idx) {
    const uint32_t res = or_terrain_res(target);
    const float step = size / (float)(res - 1);
    const float half = size * 0.5f;
    for (uint32_t y = 0; y < res; ++y) {
        for (uint32_t x = 0; x < res; ++x) {
            const uint32_t v = y * res + x;
            float px, pz, py;
            if (fused) {
                px = std::fmaf((float)x, step, -half);
                pz = std::fmaf((float)y, step, -half);
                float fx = px * 0.05f, fz = pz * 0.05f;
                float arg = std::fmaf(fx, 2.5f, fz * 1.7f);
                py = std::fmaf(std::sin(fx) * std::cos(fz), 5.0f, std::sin(arg) * 2.0f);
            } else {
                px = -half + (float)x * step;
                pz = -half + (float)y * step;
                float fx = px * 0.05f, fz = pz * 0.05f;
                float a = std::sin(fx) * std::cos(fz) * 5.0f;
                float b = std::sin(fx * 2.5f + fz * 1.7f) * 2.0f;
                py = a + b;
            }
            pos[v * 3 + 0] = px;
            pos[v * 3 + 1] = py;
            pos[v * 3 + 2] = pz;
        }
    }
    uint32_t k = 0;
    for (uint32_t y = 0; y + 1 < res; ++y) {
        for (uint32_t x = 0; x + 1 < res; ++x) {
            uint32_t i0 = y * res + x, i1 = y * res + x + 1;
            uint32_t i2 = (y + 1) * res + x, i3 = (y + 1) * res + x + 1;
            idx[k++] = i0; idx[k++] = i2; idx[k++] = i1;
            idx[k++] = i1; idx[k++] = i2; idx[k++] = i3;
        }
    }
}

void or_terrain(uint32_t target, float size, float* pos, uint32_t* idx) {
    or_terrain_ex(target, size, 0, pos, idx);
}

// include/radv_bvh.hpp:224-283 generate_dynamic_grid_mesh: unindexed grid,
// z = 5 + 0.2 sin(2x+t) cos(2y+t), 9 floats per triangle + analytic centroid.
void or_dynamic_grid(uint32_t n, float t, float extent, float* xyz9, float* cen) {
    uint32_t quads = n / 2;
    uint32_t cols = (uint32_t)std::sqrt((double)quads);
    while (cols > 1 && quads % cols != 0) cols--;
    if (cols == 0) return;
    uint32_t rows = quads / cols;
    float dx = (2.0f * extent) / (float)cols;
    float dy = (2.0f * extent) / (float)rows;
    auto z = [t](float x, float y) {
        return 5.0f + 0.2f * std::sin(x * 2.0f + t) * std::cos(y * 2.0f + t);
    };
    uint32_t tri = 0;
    for (uint32_t r = 0; r < rows; ++r) {
        for (uint32_t c = 0; c < cols; ++c) {
            float x0 = -extent + (float)c * dx, x1 = x0 + dx;
            float y0 = -extent + (float)r * dy, y1 = y0 + dy;
            float z00 = z(x0, y0), z10 = z(x1, y0), z01 = z(x0, y1), z11 = z(x1, y1);
            const float t1[9] = {x0, y0, z00, x1, y0, z10, x0, y1, z01};
            const float t2[9] = {x1, y0, z10, x1, y1, z11, x0, y1, z01};
            std::memcpy(xyz9 + tri * 9, t1, sizeof t1);
            if (cen) {
                cen[tri * 3 + 0] = (x0 + x1 + x0) / 3.0f;
                cen[tri * 3 + 1] = (y0 + y0 + y1) / 3.0f;
                cen[tri * 3 + 2] = (z00 + z10 + z01) / 3.0f;
            }
            tri++;
            std::memcpy(xyz9 + tri * 9, t2, sizeof t2);
            if (cen) {
                cen[tri * 3 + 0] = (x1 + x1 + x0) / 3.0f;
                cen[tri * 3 + 1] = (y0 + y1 + y1) / 3.0f;
                cen[tri * 3 + 2] = (z10 + z11 + z01) / 3.0f;
            }
            tri++;
        }
    }
}

// Indexed wave grid over [-e,e]^2 with the radv_bvh.hpp:257-279 triangle split
// (tri 1 = (x0,y0),(x1,y0),(x0,y1); tri 2 = (x1,y0),(x1,y1),(x0,y1)) and
// z = amp * sin(fx*x + t) * cos(fy*y).  The 64x64, amp 0.1, f 4 instance is the
// small closest-hit workload; the product generates the same mesh on device.
void or_wave_grid(uint32_t nx, uint32_t ny, float e, float t, float amp, float fx,
                  float fy, float* pos, uint32_t* idx) {
    const float dx = (2.0f * e) / (float)nx;
    const float dy = (2.0f * e) / (float)ny;
    for (uint32_t r = 0; r <= ny; ++r) {
        for (uint32_t c = 0; c <= nx; ++c) {
            uint32_t v = r * (nx + 1) + c;
            float x = -e + (float)c * dx;
            float y = -e + (float)r * dy;
            pos[v * 3 + 0] = x;
            pos[v * 3 + 1] = y;
            pos[v * 3 + 2] = amp * std::sin(fx * x + t) * std::cos(fy * y);
        }
    }
    uint32_t k = 0;
    for (uint32_t r = 0; r < ny; ++r) {
        for (uint32_t c = 0; c < nx; ++c) {
            uint32_t v00 = r * (nx + 1) + c, v10 = v00 + 1;
            uint32_t v01 = v00 + (nx + 1), v11 = v01 + 1;
            idx[k++] = v00; idx[k++] = v10; idx[k++] = v01;
            idx[k++] = v10; idx[k++] = v11; idx[k++] = v01;
        }
    }
}

// UV sphere of nu x nv quads (2*nu*nv triangles), vertex (i,j) at
// theta = pi*j/nv, phi = 2*pi*i/nu, radius r*(1 + disp*(2u-1)) with u a seeded
// hash of the vertex (pole rows share one radius so pole triangles collapse).
void or_sphere(uint32_t nu, uint32_t nv, float radius, float disp, uint32_t seed,
               float* pos, uint32_t* idx) {
    const float kPi = 3.14159265358979f;
    for (uint32_t j = 0; j <= nv; ++j) {
        float theta = kPi * (float)j / (float)nv;
        float st = std::sin(theta), ct = std::cos(theta);
        for (uint32_t i = 0; i < nu; ++i) {
            uint32_t v = j * nu + i;
            uint32_t hv = (j == 0 || j == nv) ? j * nu : v;
            float phi = 2.0f * kPi * (float)i / (float)nu;
            float r = radius * (1.0f + disp * (2.0f * hash01(seed, hv) - 1.0f));
            pos[v * 3 + 0] = r * (st * std::cos(phi));
            pos[v * 3 + 1] = r * (st * std::sin(phi));
            pos[v * 3 + 2] = r * ct;
        }
    }
    uint32_t k = 0;
    for (uint32_t j = 0; j < nv; ++j) {
        for (uint32_t i = 0; i < nu; ++i) {
            uint32_t i1 = (i + 1 == nu) ? 0 : i + 1;
            uint32_t a = j * nu + i, b = j * nu + i1;
            uint32_t c = (j + 1) * nu + i, d = (j + 1) * nu + i1;
            idx[k++] = a; idx[k++] = c; idx[k++] = b;
            idx[k++] = b; idx[k++] = c; idx[k++] = d;
        }
    }
}

// Seeded random-displaced heightfield: nx x ny quads over [-e,e]^2,
// z = amp * (2u - 1) per vertex, same split as or_wave_grid.
void or_heightfield(uint32_t nx, uint32_t ny, float e, float amp, uint32_t seed,
                    float* pos, uint32_t* idx) {
    const float dx = (2.0f * e) / (float)nx;
    const float dy = (2.0f * e) / (float)ny;
    for (uint32_t r = 0; r <= ny; ++r) {
        for (uint32_t c = 0; c <= nx; ++c) {
            uint32_t v = r * (nx + 1) + c;
            pos[v * 3 + 0] = -e + (float)c * dx;
            pos[v * 3 + 1] = -e + (float)r * dy;
            pos[v * 3 + 2] = amp * (2.0f * hash01(seed, v) - 1.0f);
        }
    }
    uint32_t k = 0;
    for (uint32_t r = 0; r < ny; ++r) {
        for (uint32_t c = 0; c < nx; ++c) {
            uint32_t v00 = r * (nx + 1) + c, v10 = v00 + 1;
            uint32_t v01 = v00 + (nx + 1), v11 = v01 + 1;
            idx[k++] = v00; idx[k++] = v10; idx[k++] = v01;
            idx[k++] = v10; idx[k++] = v11; idx[k++] = v01;
        }
    }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// software_bvh builder (common/software_bvh.hpp:24-176)
// ---------------------------------------------------------------------------
namespace {

// software_bvh.hpp:28-33 -- same member order, so std::nth_element sees the
// same sequence of comparisons and moves as the reference.
struct BuildPrimitive {
    uint32_t index;
    float center[3];
    float min[3];
    float max[3];
};

struct Bvh2Builder {
    std::vector<or_bvh2_node> nodes;
    std::vector<BuildPrimitive> ordered;

    // software_bvh.hpp:90-175 recursiveBuild
    int32_t build(std::vector<BuildPrimitive>& prims, uint32_t start, uint32_t end,
                  uint32_t maxPrims) {
        int32_t nodeIdx = (int32_t)nodes.size();
        nodes.emplace_back();
        float bmin[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
        float bmax[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (uint32_t i = start; i < end; ++i)
            for (int c = 0; c < 3; ++c) {
                bmin[c] = std::min(bmin[c], prims[i].min[c]);
                bmax[c] = std::max(bmax[c], prims[i].max[c]);
            }
        for (int c = 0; c < 3; ++c) {
            nodes[nodeIdx].aabbMin[c] = bmin[c];
            nodes[nodeIdx].aabbMax[c] = bmax[c];
        }
        const uint32_t count = end - start;
        auto make_leaf = [&]() {
            uint32_t first = (uint32_t)ordered.size();
            for (uint32_t i = start; i < end; ++i) ordered.push_back(prims[i]);
            nodes[nodeIdx].leftChild = (int32_t)~first;
            nodes[nodeIdx].rightChild = (int32_t)count;
            return nodeIdx;
        };
        if (count <= maxPrims) return make_leaf();  // :116-125

        float cmin[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
        float cmax[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (uint32_t i = start; i < end; ++i)
            for (int c = 0; c < 3; ++c) {
                cmin[c] = std::min(cmin[c], prims[i].center[c]);
                cmax[c] = std::max(cmax[c], prims[i].center[c]);
            }
        int axis = 0;  // :136-145 longest centroid axis, ties keep the lower axis
        float ext = cmax[0] - cmin[0];
        if (cmax[1] - cmin[1] > ext) { axis = 1; ext = cmax[1] - cmin[1]; }
        if (cmax[2] - cmin[2] > ext) { axis = 2; ext = cmax[2] - cmin[2]; }
        if (ext < 1e-6f) return make_leaf();  // :147-156

        uint32_t mid = (start + end) / 2;  // :159-167 median split
        std::nth_element(prims.begin() + start, prims.begin() + mid, prims.begin() + end,
                         [axis](const BuildPrimitive& a, const BuildPrimitive& b) {
                             return a.center[axis] < b.center[axis];
                         });
        int32_t l = build(prims, start, mid, maxPrims);
        int32_t r = build(prims, mid, end, maxPrims);
        nodes[nodeIdx].leftChild = l;
        nodes[nodeIdx].rightChild = r;
        return nodeIdx;
    }
};

// software_bvh.hpp:202-291 collapse (greedy largest-surface-area expansion)
int32_t collapse(const or_bvh2_node* n2, int32_t idx, std::vector<or_bvh4_node>& out) {
    const or_bvh2_node& node = n2[idx];
    if (node.leftChild < 0) {  // :205-219 leaf root -> one-child node
        int32_t o = (int32_t)out.size();
        out.emplace_back();
        or_bvh4_node& n4 = out[o];
        std::memset(&n4, 0, sizeof n4);
        n4.minX[0] = node.aabbMin[0]; n4.maxX[0] = node.aabbMax[0];
        n4.minY[0] = node.aabbMin[1]; n4.maxY[0] = node.aabbMax[1];
        n4.minZ[0] = node.aabbMin[2]; n4.maxZ[0] = node.aabbMax[2];
        n4.child[0] = node.leftChild;
        n4.count[0] = node.rightChild;
        return o;
    }
    int32_t cand[4] = {node.leftChild, node.rightChild, 0, 0};
    size_t nc = 2;
    while (nc < 4) {  // :225-247
        int best = -1;
        float bestSa = -1.0f;
        for (size_t i = 0; i < nc; ++i) {
            const or_bvh2_node& c = n2[cand[i]];
            if (c.leftChild >= 0) {
                float dx = c.aabbMax[0] - c.aabbMin[0];
                float dy = c.aabbMax[1] - c.aabbMin[1];
                float dz = c.aabbMax[2] - c.aabbMin[2];
                float sa = 2.0f * (dx * dy + dy * dz + dz * dx);
                if (sa > bestSa) { bestSa = sa; best = (int)i; }
            }
        }
        if (best == -1) break;
        int32_t e = cand[best];
        cand[best] = n2[e].leftChild;
        cand[nc++] = n2[e].rightChild;
    }
    int32_t o = (int32_t)out.size();
    out.emplace_back();
    or_bvh4_node tmp;
    std::memset(&tmp, 0, sizeof tmp);
    for (size_t i = 0; i < nc; ++i) {  // :256-269
        const or_bvh2_node& c = n2[cand[i]];
        tmp.minX[i] = c.aabbMin[0]; tmp.maxX[i] = c.aabbMax[0];
        tmp.minY[i] = c.aabbMin[1]; tmp.maxY[i] = c.aabbMax[1];
        tmp.minZ[i] = c.aabbMin[2]; tmp.maxZ[i] = c.aabbMax[2];
        if (c.leftChild < 0) {
            tmp.child[i] = c.leftChild;
            tmp.count[i] = c.rightChild;
        } else {
            tmp.count[i] = -1;
        }
    }
    for (size_t i = 0; i < nc; ++i)  // :271-277 recurse after reserving the node
        if (n2[cand[i]].leftChild >= 0) tmp.child[i] = collapse(n2, cand[i], out);
    out[o] = tmp;
    return o;
}

// Treelet restructuring of a BVH2 whose leaves are triangle ranges (Karras &
// Aila 2013; no reference counterpart), one round = a post-order pass.  At
// inner node x: a treelet of up to 7 leaves grows from x's children by
// expanding the inner treelet leaf of largest surface area (first in slot
// order on ties); over its leaf subsets s (ascending), cost(s) = 1.2 * SA(box
// of s) + the cheapest split of s (the part holding s's lowest leaf, scanned
// from s - 1 down, first strict minimum), a single leaf costing its subtree's
// SAH cost (leaf: SA * triangles; inner: 1.2 * SA + both children's costs).
// If cost(all) is strictly below x's current cost, the treelet is rebuilt
// breadth-first, reusing its inner nodes in the order they were expanded; the
// rebuilt nodes' boxes are the unions of their children.
struct TreeletOpt {
    std::vector<or_bvh2_node>& t;
    std::vector<float> cost;
    static constexpr int kLeaves = 7;
    static constexpr float kInner = 1.2f;

    static float area(const float* lo, const float* hi) {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
    bool inner(int32_t u) const { return t[u].leftChild >= 0; }
    float sah(int32_t u) const {
        return inner(u) ? cost[u] : area(t[u].aabbMin, t[u].aabbMax) * (float)t[u].rightChild;
    }
    static float split(const float* C, uint32_t s, uint32_t& bp) {
        const uint32_t low = s & (0u - s);
        float best = INFINITY;
        bp = 0;
        for (uint32_t p = (s - 1) & s; p > 0; p = (p - 1) & s) {
            if (!(p & low)) continue;
            const float c = C[p] + C[s ^ p];
            if (c < best) {
                best = c;
                bp = p;
            }
        }
        return best;
    }
    void node(int32_t x) {
        int32_t lf[kLeaves], in[kLeaves - 1];
        lf[0] = t[x].leftChild;
        lf[1] = t[x].rightChild;
        in[0] = x;
        int nl = 2, ni = 1;
        float cx = kInner * area(t[x].aabbMin, t[x].aabbMax) + (sah(lf[0]) + sah(lf[1]));
        while (nl < kLeaves) {
            int best = -1;
            float bs = -1.0f;
            for (int q = 0; q < nl; ++q) {
                if (!inner(lf[q])) continue;
                const float a = area(t[lf[q]].aabbMin, t[lf[q]].aabbMax);
                if (a > bs) {
                    bs = a;
                    best = q;
                }
            }
            if (best < 0) break;
            const int32_t e = lf[best];
            in[ni++] = e;
            lf[best] = t[e].leftChild;
            lf[nl++] = t[e].rightChild;
        }
        if (nl >= 3) {
            float C[1 << kLeaves];
            for (int q = 0; q < nl; ++q) C[1u << q] = sah(lf[q]);
            const uint32_t full = (1u << nl) - 1;
            for (uint32_t s = 3; s <= full; ++s) {
                if (!(s & (s - 1))) continue;
                float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
                for (int q = 0; q < nl; ++q) {
                    if (!((s >> q) & 1u)) continue;
                    for (int a = 0; a < 3; ++a) {
                        lo[a] = std::min(lo[a], t[lf[q]].aabbMin[a]);
                        hi[a] = std::max(hi[a], t[lf[q]].aabbMax[a]);
                    }
                }
                uint32_t bp;
                const float sp = split(C, s, bp);
                C[s] = kInner * area(lo, hi) + sp;
            }
            if (C[full] < cx) {
                int32_t qn[kLeaves - 1], ql[kLeaves - 1], qr[kLeaves - 1];
                uint32_t qs[kLeaves - 1];
                qn[0] = x;
                qs[0] = full;
                int nq = 1, next = 1;
                bool ok = true;
                for (int i = 0; i < nq; ++i) {
                    uint32_t p;
                    split(C, qs[i], p);
                    if (p == 0) {
                        ok = false;
                        break;
                    }
                    const uint32_t part[2] = {p, qs[i] ^ p};
                    int32_t ch[2];
                    for (int j = 0; j < 2; ++j) {
                        if (!(part[j] & (part[j] - 1))) {
                            ch[j] = lf[__builtin_ctz(part[j])];
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
                    for (int i = 0; i < nq; ++i) {
                        t[qn[i]].leftChild = ql[i];
                        t[qn[i]].rightChild = qr[i];
                    }
                    for (int i = nq - 1; i >= 1; --i) {
                        or_bvh2_node& y = t[qn[i]];
                        for (int a = 0; a < 3; ++a) {
                            y.aabbMin[a] = std::min(t[ql[i]].aabbMin[a], t[qr[i]].aabbMin[a]);
                            y.aabbMax[a] = std::max(t[ql[i]].aabbMax[a], t[qr[i]].aabbMax[a]);
                        }
                        cost[qn[i]] = C[qs[i]];
                    }
                    cx = C[full];
                }
            }
        }
        cost[x] = cx;
    }
    void post(int32_t x) {
        if (!inner(x)) return;
        post(t[x].leftChild);
        post(t[x].rightChild);
        node(x);
    }
};

}  // namespace

extern "C" {

// software_bvh.hpp:35-87 SoftwareBVH::build.  Returns 0 on success, 1 if
// node_cap is too small.  tris48 receives GpuTriangle records (w = 0) in leaf
// order; order[i] receives the original triangle index of leaf slot i, which
// the reference discards (its GpuTriangle has no ID).
int or_bvh2_build(const float* pos, uint32_t nverts, const uint32_t* idx, uint32_t ntris,
                  uint32_t maxPrims, or_bvh2_node* nodes_out, uint32_t node_cap,
                  uint32_t* node_count, float* tris48, uint32_t* order) {
    (void)nverts;
    std::vector<BuildPrimitive> prims(ntris);
    for (uint32_t i = 0; i < ntris; ++i) {  // :40-55
        const float* p0 = pos + 3 * idx[i * 3 + 0];
        const float* p1 = pos + 3 * idx[i * 3 + 1];
        const float* p2 = pos + 3 * idx[i * 3 + 2];
        prims[i].index = i;
        for (int c = 0; c < 3; ++c) {
            prims[i].min[c] = std::min({p0[c], p1[c], p2[c]});
            prims[i].max[c] = std::max({p0[c], p1[c], p2[c]});
            prims[i].center[c] = (p0[c] + p1[c] + p2[c]) / 3.0f;
        }
    }
    Bvh2Builder b;
    b.nodes.reserve((size_t)ntris * 2);
    b.ordered.reserve(ntris);
    if (ntris == 0) {  // the reference emits one empty leaf: bounds FLT_MAX/-FLT_MAX
        b.build(prims, 0, 0, maxPrims);
    } else {
        b.build(prims, 0, ntris, maxPrims);
    }
    if (b.nodes.size() > node_cap) return 1;
    std::memcpy(nodes_out, b.nodes.data(), b.nodes.size() * sizeof(or_bvh2_node));
    *node_count = (uint32_t)b.nodes.size();
    for (uint32_t i = 0; i < ntris; ++i) {  // :65-84 leaf-order triangles
        uint32_t o = b.ordered[i].index;
        if (order) order[i] = o;
        if (tris48) {
            for (int v = 0; v < 3; ++v) {
                const float* p = pos + 3 * idx[o * 3 + v];
                tris48[i * 12 + v * 4 + 0] = p[0];
                tris48[i * 12 + v * 4 + 1] = p[1];
                tris48[i * 12 + v * 4 + 2] = p[2];
                tris48[i * 12 + v * 4 + 3] = 0.0f;
            }
        }
    }
    return 0;
}

// software_bvh.hpp:193-199 SoftwareBVH4::fromBVH2
int or_bvh4_collapse(const or_bvh2_node* n2, uint32_t count2, or_bvh4_node* n4,
                     uint32_t cap4, uint32_t* count4) {
    std::vector<or_bvh4_node> out;
    if (count2 > 0) collapse(n2, 0, out);
    if (out.size() > cap4) return 1;
    if (!out.empty()) std::memcpy(n4, out.data(), out.size() * sizeof(or_bvh4_node));
    *count4 = (uint32_t)out.size();
    return 0;
}

// GpuTriangle (48 B, software_bvh.hpp:18-22) + leaf-order IDs -> 64-B leaf
// records laid out as RadvBvhTriangleNode (radv_bvh.hpp:142-152), with the
// original triangle index in triangle_id, opaque flag, barycentrics id 9.
void or_tris48_to_tri64(const float* t48, const uint32_t* order, uint32_t n, or_tri64* out) {
    for (uint32_t i = 0; i < n; ++i) {
        or_tri64 r;
        std::memset(&r, 0, sizeof r);
        for (int v = 0; v < 3; ++v)
            for (int c = 0; c < 3; ++c) r.v[v * 3 + c] = t48[i * 12 + v * 4 + c];
        r.triangle_id = order ? order[i] : i;
        r.geometry_id_and_flags = 0x80000000u;
        r.id = 9u;
        out[i] = r;
    }
}

// Number of node levels on the longest root-to-leaf path.
uint32_t or_bvh4_depth(const or_bvh4_node* nodes, uint32_t count) {
    if (count == 0) return 0;
    std::vector<std::pair<int32_t, uint32_t>> st{{0, 1u}};
    uint32_t best = 0;
    while (!st.empty()) {
        auto [n, d] = st.back();
        st.pop_back();
        best = std::max(best, d);
        for (int k = 0; k < 4; ++k)
            if (nodes[n].count[k] < 0) st.push_back({nodes[n].child[k], d + 1});
    }
    return best;
}

// ---------------------------------------------------------------------------
// Quantized nodes (rt_qnode4): 8-bit child boxes on a per-node grid, rounded
// outward (no reference counterpart; the compressed wide BVH of Ylitie,
// Karras & Laine 2017 with four children)
// ---------------------------------------------------------------------------

static inline float q_dec(float p, int q, float s) { return p + (float)q * s; }

void or_bvh4_quantize(const or_bvh4_node* nodes, uint32_t count, or_qnode4* out) {
    for (uint32_t i = 0; i < count; ++i) {
        const or_bvh4_node& n = nodes[i];
        const float* lo[3] = {n.minX, n.minY, n.minZ};
        const float* hi[3] = {n.maxX, n.maxY, n.maxZ};
        or_qnode4 q;
        std::memset(&q, 0, sizeof q);
        for (int a = 0; a < 3; ++a) {
            float p = INFINITY, m = -INFINITY;
            for (int k = 0; k < 4; ++k)
                if (n.count[k] != 0) {
                    p = std::min(p, lo[a][k]);
                    m = std::max(m, hi[a][k]);
                }
            if (!(p <= m)) p = m = 0.0f;
            const int e = least_true(-126, 127, [&](int x) {
                return q_dec(p, 255, std::ldexp(1.0f, x)) >= m;
            });
            const float s = std::ldexp(1.0f, e);
            q.p[a] = p;
            q.e[a] = (int8_t)e;
            for (int k = 0; k < 4; ++k) {
                if (n.count[k] == 0) continue;
                // greatest code at or below the min: 255 - least code above it
                const float l = lo[a][k], h = hi[a][k];
                const int above = least_true(0, 256, [&](int c) {
                    return c > 255 || q_dec(p, c, s) > l;
                });
                q.qlo[a][k] = (uint8_t)(above > 0 ? above - 1 : 0);
                q.qhi[a][k] = (uint8_t)std::min(255, least_true(0, 255, [&](int c) {
                    return q_dec(p, c, s) >= h;
                }));
            }
        }
        for (int k = 0; k < 4; ++k) {
            q.count[k] = (int16_t)n.count[k];
            q.child[k] = n.count[k] != 0 ? n.child[k] : 0;
        }
        out[i] = q;
    }
}

void or_bvh4_dequantize(const or_qnode4* q, uint32_t count, or_bvh4_node* out) {
    for (uint32_t i = 0; i < count; ++i) {
        or_bvh4_node n;
        std::memset(&n, 0, sizeof n);
        float* lo[3] = {n.minX, n.minY, n.minZ};
        float* hi[3] = {n.maxX, n.maxY, n.maxZ};
        for (int a = 0; a < 3; ++a) {
            const float s = std::ldexp(1.0f, q[i].e[a]);
            for (int k = 0; k < 4; ++k) {
                lo[a][k] = q_dec(q[i].p[a], q[i].qlo[a][k], s);
                hi[a][k] = q_dec(q[i].p[a], q[i].qhi[a][k], s);
            }
        }
        for (int k = 0; k < 4; ++k) {
            n.child[k] = q[i].child[k];
            n.count[k] = q[i].count[k];
        }
        out[i] = n;
    }
}

// ---------------------------------------------------------------------------
// Traversal: benchmarks/exp0_probe/shaders/software_bvh_bvh4.comp:96-190
// ---------------------------------------------------------------------------

// Moller-Trumbore exactly as software_bvh_bvh4.comp:166-182 (|a| < 1e-7
// rejects, inclusive edges).  Returns true and t when the ray crosses the
// triangle's plane inside the triangle.
static inline bool mt_hit(const or_tri64& tri, const float o[3], const float d[3], float* tout) {
    float e1x = tri.v[3] - tri.v[0], e1y = tri.v[4] - tri.v[1], e1z = tri.v[5] - tri.v[2];
    float e2x = tri.v[6] - tri.v[0], e2y = tri.v[7] - tri.v[1], e2z = tri.v[8] - tri.v[2];
    float hx = d[1] * e2z - d[2] * e2y;
    float hy = d[2] * e2x - d[0] * e2z;
    float hz = d[0] * e2y - d[1] * e2x;
    float a = e1x * hx + e1y * hy + e1z * hz;
    if (std::fabs(a) < 1e-7f) return false;
    float f = 1.0f / a;
    float sx = o[0] - tri.v[0], sy = o[1] - tri.v[1], sz = o[2] - tri.v[2];
    float u = f * (sx * hx + sy * hy + sz * hz);
    if (u < 0.0f || u > 1.0f) return false;
    float qx = sy * e1z - sz * e1y;
    float qy = sz * e1x - sx * e1z;
    float qz = sx * e1y - sy * e1x;
    float v = f * (d[0] * qx + d[1] * qy + d[2] * qz);
    if (v < 0.0f || u + v > 1.0f) return false;
    *tout = f * (e2x * qx + e2y * qy + e2z * qz);
    return true;
}

// Closest-hit (mode 0) or any-hit (mode 1) BVH4 traversal.  Node order, child
// sort and pruning follow software_bvh_bvh4.comp:111-190 in both modes; any
// hit (the TerminateOnFirstHit queries of rq_ao.comp:78-90 /
// tiled_lighting_ra.comp:138-150) stops at the first accepted triangle
// (near-first order finds occluders of surface-bound shadow / AO rays after
// ~40 % fewer node visits than slot order).  Ray tmin/tmax come from the ray
// (the shader hard-codes 0.01 / 1000).  stack_limit 24
// reproduces the shader's silent drop on overflow (:185-187); 0 = unbounded.
// out_prim = triangle_id of the accepted leaf record (0xFFFFFFFF on a miss),
// out_t = hit t (tmax on a miss).
void or_trace_bvh4(const or_bvh4_node* nodes, uint32_t node_count, const or_tri64* tris,
                   const float* ox, const float* oy, const float* oz, const float* tmn,
                   const float* dx, const float* dy, const float* dz, const float* tmx,
                   uint32_t nrays, int mode, uint32_t stack_limit, int nthreads,
                   uint32_t* out_prim, float* out_t, uint32_t* out_box, uint32_t* out_tri) {
    parallel_chunks(nrays, nthreads, [&](uint32_t b, uint32_t e) {
        std::vector<int32_t> stack(1024);
        for (uint32_t r = b; r < e; ++r) {
            const float o[3] = {ox[r], oy[r], oz[r]};
            const float d[3] = {dx[r], dy[r], dz[r]};
            float inv[3];
            for (int c = 0; c < 3; ++c)  // :96-100
                inv[c] = std::fabs(d[c]) > 1e-6f ? 1.0f / d[c] : 1e6f;
            const float rayTMin = tmn[r];
            float hitT = tmx[r];
            uint32_t hitPrim = 0xFFFFFFFFu;
            uint32_t boxTests = 0, triTests = 0;
            uint32_t ptr = 0;
            bool done = false;
            if (node_count > 0) stack[ptr++] = 0;
            while (ptr > 0 && !done) {
                const or_bvh4_node& n = nodes[stack[--ptr]];
                boxTests += 4;
                float tEntry[4], tExit[4];
                for (int k = 0; k < 4; ++k) {  // :116-132 slab test
                    float t0x = (n.minX[k] - o[0]) * inv[0], t1x = (n.maxX[k] - o[0]) * inv[0];
                    float t0y = (n.minY[k] - o[1]) * inv[1], t1y = (n.maxY[k] - o[1]) * inv[1];
                    float t0z = (n.minZ[k] - o[2]) * inv[2], t1z = (n.maxZ[k] - o[2]) * inv[2];
                    float tminX = std::min(t0x, t1x), tmaxX = std::max(t0x, t1x);
                    float tminY = std::min(t0y, t1y), tmaxY = std::max(t0y, t1y);
                    float tminZ = std::min(t0z, t1z), tmaxZ = std::max(t0z, t1z);
                    tEntry[k] = std::max(std::max(tminX, tminY), std::max(tminZ, rayTMin));
                    tExit[k] = std::min(std::min(tmaxX, tmaxY), std::min(tmaxZ, hitT));
                }
                int slot[4];
                float st[4];
                int hc = 0;
                for (int k = 0; k < 4; ++k)  // :138-144
                    if (n.count[k] != 0 && tExit[k] >= tEntry[k]) {
                        slot[hc] = k;
                        st[hc] = tEntry[k];
                        hc++;
                    }
                for (int i = 0; i < hc - 1; ++i)  // :149-156 descending by entry t
                    for (int j = i + 1; j < hc; ++j)
                        if (st[i] < st[j]) {
                            std::swap(st[i], st[j]);
                            std::swap(slot[i], slot[j]);
                        }
                for (int i = 0; i < hc && !done; ++i) {  // :158-189
                    int k = slot[i];
                    if (n.count[k] > 0) {
                        uint32_t first = (uint32_t)~n.child[k];
                        uint32_t cnt = (uint32_t)n.count[k];
                        triTests += cnt;
                        for (uint32_t p = 0; p < cnt; ++p) {
                            float t;
                            if (!mt_hit(tris[first + p], o, d, &t)) continue;
                            if (t >= rayTMin && t < hitT) {
                                hitT = t;
                                hitPrim = tris[first + p].triangle_id;
                                if (mode == OR_ANY) { done = true; break; }
                            }
                        }
                    } else if (n.count[k] < 0) {
                        if (stack_limit == 0 || ptr < stack_limit) stack[ptr++] = n.child[k];
                    }
                }
            }
            out_prim[r] = hitPrim;
            if (out_t) out_t[r] = hitT;
            if (out_box) out_box[r] = boxTests;
            if (out_tri) out_tri[r] = triTests;
        }
    });
}

// software_bvh_procedural.comp:47-82 intersectAABB: slabs against the ray,
// hit iff tFar >= max(0, tNear) and tNear < tCurrent (the ray's tmin is not
// used by the box test).
static inline bool bvh2_box(const float bmin[3], const float bmax[3], const float o[3],
                            const float inv[3], float tcur, float* tbox) {
    float tn = -FLT_MAX, tf = FLT_MAX;
    float t0[3], t1[3];
    for (int c = 0; c < 3; ++c) {
        t0[c] = (bmin[c] - o[c]) * inv[c];
        t1[c] = (bmax[c] - o[c]) * inv[c];
    }
    tn = std::max(std::max(std::min(t0[0], t1[0]), std::min(t0[1], t1[1])),
                  std::min(t0[2], t1[2]));
    tf = std::min(std::min(std::max(t0[0], t1[0]), std::max(t0[1], t1[1])),
                  std::max(t0[2], t1[2]));
    *tbox = tn;
    return tf >= std::max(0.0f, tn) && tn < tcur;
}

// software_bvh_procedural.comp:60-82 intersectTriangle on a 48-B GpuTriangle
// (v0, v1, v2 as float4): Moller-Trumbore with |a| < 1e-8 rejected.
static inline bool bvh2_tri(const float* t48, const float o[3], const float d[3], float* tout) {
    const float *v0 = t48, *v1 = t48 + 4, *v2 = t48 + 8;
    float e1x = v1[0] - v0[0], e1y = v1[1] - v0[1], e1z = v1[2] - v0[2];
    float e2x = v2[0] - v0[0], e2y = v2[1] - v0[1], e2z = v2[2] - v0[2];
    float hx = d[1] * e2z - d[2] * e2y;
    float hy = d[2] * e2x - d[0] * e2z;
    float hz = d[0] * e2y - d[1] * e2x;
    float a = e1x * hx + e1y * hy + e1z * hz;
    if (std::fabs(a) < 1e-8f) return false;
    float f = 1.0f / a;
    float sx = o[0] - v0[0], sy = o[1] - v0[1], sz = o[2] - v0[2];
    float u = f * (sx * hx + sy * hy + sz * hz);
    if (u < 0.0f || u > 1.0f) return false;
    float qx = sy * e1z - sz * e1y;
    float qy = sz * e1x - sx * e1z;
    float qz = sx * e1y - sy * e1x;
    float v = f * (d[0] * qx + d[1] * qy + d[2] * qz);
    if (v < 0.0f || u + v > 1.0f) return false;
    *tout = f * (e2x * qx + e2y * qy + e2z * qz);
    return true;
}

// BVH2 traversal of a SoftwareBVH (GpuBvhNode + GpuTriangle arrays), the
// loop of software_bvh_procedural.comp:119-171: the popped node's box is
// tested (culled when it misses or starts beyond the closest t), a leaf
// tests its triangles in record order (accept t > tmin and t < closest t:
// the shader's t > 0.01 with the ray's tmin), an inner node tests both
// children's boxes and pushes the hit ones far first (left popped first
// when tLeft < tRight).  inv = 1 / d with |d| <= 1e-8 replaced by +-1e-8
// (:108-110).  The shader starts from closestT = 1000 and writes a hit flag;
// here closest t starts at the ray's tmax and the hit's leaf position (or
// order[position], the original index) and t are written.  Any hit (mode 1)
// stops at the first accepted triangle.  Counters: box tests = boxes tested
// (popped nodes + children of inner nodes), triangle tests = triangles
// tested.  The stack is unbounded (the shader's holds 32).
void or_trace_bvh2(const or_bvh2_node* nodes, uint32_t node_count, const float* tris48,
                   const uint32_t* order, const float* ox, const float* oy, const float* oz,
                   const float* tmn, const float* dx, const float* dy, const float* dz,
                   const float* tmx, uint32_t nrays, int mode, int nthreads,
                   uint32_t* out_prim, float* out_t, uint32_t* out_box, uint32_t* out_tri) {
    parallel_chunks(nrays, nthreads, [&](uint32_t b, uint32_t e) {
        std::vector<uint32_t> stack(256);
        for (uint32_t r = b; r < e; ++r) {
            const float o[3] = {ox[r], oy[r], oz[r]};
            const float d[3] = {dx[r], dy[r], dz[r]};
            float inv[3];
            for (int c = 0; c < 3; ++c) {
                const float dd = std::fabs(d[c]) > 1e-8f ? d[c] : (d[c] >= 0.0f ? 1e-8f : -1e-8f);
                inv[c] = 1.0f / dd;
            }
            float closest = tmx[r];
            uint32_t prim = 0xFFFFFFFFu, boxT = 0, triT = 0;
            size_t sp = 0;
            bool done = false;
            if (node_count > 0) stack[sp++] = 0;
            while (sp > 0 && !done) {
                const or_bvh2_node& nd = nodes[stack[--sp]];
                float tb;
                ++boxT;
                if (!bvh2_box(nd.aabbMin, nd.aabbMax, o, inv, closest, &tb)) continue;
                if (nd.leftChild < 0) {
                    const uint32_t first = (uint32_t)~nd.leftChild, cnt = (uint32_t)nd.rightChild;
                    for (uint32_t i = 0; i < cnt; ++i) {
                        ++triT;
                        float t;
                        if (!bvh2_tri(tris48 + 12 * (size_t)(first + i), o, d, &t)) continue;
                        if (t > tmn[r] && t < closest) {
                            closest = t;
                            prim = order ? order[first + i] : first + i;
                            if (mode == OR_ANY) { done = true; break; }
                        }
                    }
                } else {
                    const uint32_t L = (uint32_t)nd.leftChild, R = (uint32_t)nd.rightChild;
                    float tl, tr;
                    boxT += 2;
                    const bool hl = bvh2_box(nodes[L].aabbMin, nodes[L].aabbMax, o, inv, closest, &tl);
                    const bool hr = bvh2_box(nodes[R].aabbMin, nodes[R].aabbMax, o, inv, closest, &tr);
                    if (sp + 2 > stack.size()) stack.resize(stack.size() * 2);
                    if (hl && hr) {
                        if (tl < tr) { stack[sp++] = R; stack[sp++] = L; }
                        else { stack[sp++] = L; stack[sp++] = R; }
                    } else if (hl) {
                        stack[sp++] = L;
                    } else if (hr) {
                        stack[sp++] = R;
                    }
                }
            }
            out_prim[r] = prim;
            if (out_t) out_t[r] = closest;
            if (out_box) out_box[r] = boxT;
            if (out_tri) out_tri[r] = triT;
        }
    });
}

// Every triangle against every ray with the same predicate.  Closest: the
// smallest t; at equal t the lowest triangle_id.  Any: first triangle whose t
// lies in [tmin, tmax).
void or_trace_brute(const or_tri64* tris, uint32_t ntris,
                    const float* ox, const float* oy, const float* oz, const float* tmn,
                    const float* dx, const float* dy, const float* dz, const float* tmx,
                    uint32_t nrays, int mode, int nthreads, uint32_t* out_prim, float* out_t) {
    parallel_chunks(nrays, nthreads, [&](uint32_t b, uint32_t e) {
        for (uint32_t r = b; r < e; ++r) {
            const float o[3] = {ox[r], oy[r], oz[r]};
            const float d[3] = {dx[r], dy[r], dz[r]};
            float best = tmx[r];
            uint32_t prim = 0xFFFFFFFFu;
            for (uint32_t i = 0; i < ntris; ++i) {
                float t;
                if (!mt_hit(tris[i], o, d, &t)) continue;
                if (t < tmn[r] || t >= tmx[r]) continue;
                if (t < best || (t == best && tris[i].triangle_id < prim)) {
                    best = t;
                    prim = tris[i].triangle_id;
                    if (mode == OR_ANY) break;
                }
            }
            out_prim[r] = prim;
            if (out_t) out_t[r] = best;
        }
    });
}

// ---------------------------------------------------------------------------
// Ray generators
// ---------------------------------------------------------------------------

// software_bvh_bvh4.comp:50-54 pcg_hash
uint32_t or_pcg_hash(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

static inline void normalize3(float& x, float& y, float& z) {
    float l = std::sqrt(x * x + y * y + z * z);
    x = x / l; y = y / l; z = z / l;
}

// software_bvh_bvh4.comp:69-94 procedural rays: coherent 1024-wide camera grid
// (dir = normalize(fwd + 0.55 u right + 0.55 v up)); incoherent hashed origins
// in [-40,40]x[15,30]x[-40,40] aiming down.  hashFloat scales by 2^-32.
void or_exp0_rays(uint32_t count, int coherent, const float cam[3], const float look[3],
                  float* ox, float* oy, float* oz, float* dx, float* dy, float* dz) {
    const float hs = 1.0f / 4294967295.0f;
    auto hf = [hs](uint32_t s) { return (float)or_pcg_hash(s) * hs; };
    float fx = look[0] - cam[0], fy = look[1] - cam[1], fz = look[2] - cam[2];
    normalize3(fx, fy, fz);
    float rx = fy * 0.0f - fz * 1.0f, ry = fz * 0.0f - fx * 0.0f, rz = fx * 1.0f - fy * 0.0f;
    normalize3(rx, ry, rz);
    float ux = ry * fz - rz * fy, uy = rz * fx - rx * fz, uz = rx * fy - ry * fx;
    const uint32_t width = 1024, height = (count + width - 1) / width;
    for (uint32_t i = 0; i < count; ++i) {
        float o0, o1, o2, d0, d1, d2;
        if (coherent) {
            uint32_t x = i % width, y = i / width;
            float u = ((float)x + 0.5f) / (float)width * 2.0f - 1.0f;
            float v = ((float)y + 0.5f) / (float)height * 2.0f - 1.0f;
            o0 = cam[0]; o1 = cam[1]; o2 = cam[2];
            d0 = fx + rx * u * 0.55f + ux * v * 0.55f;
            d1 = fy + ry * u * 0.55f + uy * v * 0.55f;
            d2 = fz + rz * u * 0.55f + uz * v * 0.55f;
        } else {
            o0 = hf(i * 6 + 0) * 80.0f - 40.0f;
            o1 = hf(i * 6 + 1) * 15.0f + 15.0f;
            o2 = hf(i * 6 + 2) * 80.0f - 40.0f;
            d0 = hf(i * 6 + 3) * 0.8f - 0.4f;
            d1 = -1.0f - hf(i * 6 + 4) * 0.5f;
            d2 = hf(i * 6 + 5) * 0.8f - 0.4f;
        }
        normalize3(d0, d1, d2);
        ox[i] = o0; oy[i] = o1; oz[i] = o2;
        dx[i] = d0; dy[i] = d1; dz[i] = d2;
    }
}

// Pinhole camera, one ray per pixel, row 0 at the top.  The basis comes from
// eye/look/up in double precision rounded to float (the product derives it
// the same way on the host); per pixel
// dir = normalize(fwd + (sx * tan_h * aspect) right + (sy * tan_h) up).
void or_camera_rays(uint32_t w, uint32_t h, const float eye[3], const float look[3],
                    const float up[3], float vfov_deg, float tmin, float tmax,
                    float* ox, float* oy, float* oz, float* tn,
                    float* dx, float* dy, float* dz, float* tx) {
    double f[3] = {(double)look[0] - eye[0], (double)look[1] - eye[1], (double)look[2] - eye[2]};
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& c : f) c /= fl;
    double r[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    double rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (double& c : r) c /= rl;
    double u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
    double th = std::tan(vfov_deg * 3.14159265358979323846 / 360.0);
    double aspect = (double)w / (double)h;
    float F[3], R[3], U[3];
    for (int c = 0; c < 3; ++c) {
        F[c] = (float)f[c];
        R[c] = (float)(r[c] * th * aspect);
        U[c] = (float)(u[c] * th);
    }
    const float sxs = 2.0f / (float)w, sys = 2.0f / (float)h;
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            uint32_t i = y * w + x;
            float sx = ((float)x + 0.5f) * sxs - 1.0f;
            float sy = 1.0f - ((float)y + 0.5f) * sys;
            float d0 = F[0] + sx * R[0] + sy * U[0];
            float d1 = F[1] + sx * R[1] + sy * U[1];
            float d2 = F[2] + sx * R[2] + sy * U[2];
            normalize3(d0, d1, d2);
            ox[i] = eye[0]; oy[i] = eye[1]; oz[i] = eye[2]; tn[i] = tmin;
            dx[i] = d0; dy[i] = d1; dz[i] = d2; tx[i] = tmax;
        }
}

// Geometric normal of triangle prim facing against the incoming direction.
// Unit normal of triangle prim (indexed, or 9 floats per triangle when idx
// is null), turned to face against d.
static inline void face_normal(const float* pos, const uint32_t* idx, uint32_t prim,
                               float ddx, float ddy, float ddz, float n[3]) {
    const float* p0 = idx ? pos + 3 * (size_t)idx[prim * 3 + 0] : pos + 9 * (size_t)prim;
    const float* p1 = idx ? pos + 3 * (size_t)idx[prim * 3 + 1] : p0 + 3;
    const float* p2 = idx ? pos + 3 * (size_t)idx[prim * 3 + 2] : p0 + 6;
    float e1x = p1[0] - p0[0], e1y = p1[1] - p0[1], e1z = p1[2] - p0[2];
    float e2x = p2[0] - p0[0], e2y = p2[1] - p0[1], e2z = p2[2] - p0[2];
    float nx = e1y * e2z - e1z * e2y;
    float ny = e1z * e2x - e1x * e2z;
    float nz = e1x * e2y - e1y * e2x;
    normalize3(nx, ny, nz);
    if (nx * ddx + ny * ddy + nz * ddz > 0.0f) { nx = -nx; ny = -ny; nz = -nz; }
    n[0] = nx; n[1] = ny; n[2] = nz;
}

// Shadow rays in the role of tiled_lighting_ra.comp:136-155 /
// lighting_rayquery_tiled.comp:110-129: from each primary hit point
// P = o + t d, origin P + eps N, towards each light, tmin eps, tmax = distance
// from that origin to the light.  Light-major layout: ray = l * npix + p.
// A pixel whose primary ray missed gets an empty interval (tmin 1, tmax 0).
void or_shadow_rays(const float* pos, const uint32_t* idx,
                    const float* pox, const float* poy, const float* poz,
                    const float* pdx, const float* pdy, const float* pdz,
                    const uint32_t* prim, const float* t, uint32_t npix,
                    const float* lights, uint32_t nl, float eps,
                    float* ox, float* oy, float* oz, float* tn,
                    float* dx, float* dy, float* dz, float* tx) {
    for (uint32_t p = 0; p < npix; ++p) {
        if (prim[p] == 0xFFFFFFFFu) {
            for (uint32_t l = 0; l < nl; ++l) {
                uint64_t r = (uint64_t)l * npix + p;
                ox[r] = 0.0f; oy[r] = 0.0f; oz[r] = 0.0f; tn[r] = 1.0f;
                dx[r] = 0.0f; dy[r] = 0.0f; dz[r] = 1.0f; tx[r] = 0.0f;
            }
            continue;
        }
        float n[3];
        face_normal(pos, idx, prim[p], pdx[p], pdy[p], pdz[p], n);
        float Px = pox[p] + t[p] * pdx[p];
        float Py = poy[p] + t[p] * pdy[p];
        float Pz = poz[p] + t[p] * pdz[p];
        float Ox = Px + eps * n[0], Oy = Py + eps * n[1], Oz = Pz + eps * n[2];
        for (uint32_t l = 0; l < nl; ++l) {
            uint64_t r = (uint64_t)l * npix + p;
            float Lx = lights[l * 3 + 0] - Ox;
            float Ly = lights[l * 3 + 1] - Oy;
            float Lz = lights[l * 3 + 2] - Oz;
            float dist = std::sqrt(Lx * Lx + Ly * Ly + Lz * Lz);
            ox[r] = Ox; oy[r] = Oy; oz[r] = Oz; tn[r] = eps;
            dx[r] = Lx / dist; dy[r] = Ly / dist; dz[r] = Lz / dist; tx[r] = dist;
        }
    }
}

// shaders/rq_ao.comp:18-45 hash / rand / cosine hemisphere sample.  The
// hash is Jenkins' one-at-a-time hash of the single key value x (one mixing
// step, then the final avalanche).
static inline uint32_t ao_hash(uint32_t x) {
    x += (x << 10u); x ^= (x >> 6u); x += (x << 3u); x ^= (x >> 11u); x += (x << 15u);
    return x;
}
uint32_t or_ao_hash(uint32_t x) { return ao_hash(x); }

// One cosine-weighted AO ray for pixel p about the unit normal n
// (rq_ao.comp:32-97, 70-87): seed y*W + x + 5544 + frame*W*H, origin
// P + offset n, [tmin, tmax).  Missed pixels get an empty interval.
static void ao_ray(uint32_t p, const float n[3], float Px, float Py, float Pz, uint32_t w,
                   uint32_t h, uint32_t frame, float offset, float tmin, float tmax,
                   float* ox, float* oy, float* oz, float* tn,
                   float* dx, float* dy, float* dz, float* tx) {
    uint32_t seed = p + 5544u + frame * w * h;
    seed = ao_hash(seed);
    float u1 = (float)seed / 4294967295.0f;
    seed = ao_hash(seed);
    float u2 = (float)seed / 4294967295.0f;
    float r = std::sqrt(u1);
    float th = 6.2831853f * u2;
    float lx = r * std::cos(th), ly = r * std::sin(th);
    float lz = std::sqrt(std::max(0.0f, 1.0f - u1));
    float upx = 0.0f, upy = 0.0f, upz = 1.0f;
    if (!(std::fabs(n[2]) < 0.999f)) { upx = 1.0f; upz = 0.0f; }
    float tx0 = upy * n[2] - upz * n[1];
    float ty0 = upz * n[0] - upx * n[2];
    float tz0 = upx * n[1] - upy * n[0];
    normalize3(tx0, ty0, tz0);
    float bx = n[1] * tz0 - n[2] * ty0;
    float by = n[2] * tx0 - n[0] * tz0;
    float bz = n[0] * ty0 - n[1] * tx0;
    float d0 = tx0 * lx + bx * ly + n[0] * lz;
    float d1 = ty0 * lx + by * ly + n[1] * lz;
    float d2 = tz0 * lx + bz * ly + n[2] * lz;
    normalize3(d0, d1, d2);
    ox[p] = Px + offset * n[0]; oy[p] = Py + offset * n[1]; oz[p] = Pz + offset * n[2];
    tn[p] = tmin; tx[p] = tmax;
    dx[p] = d0; dy[p] = d1; dz[p] = d2;
}

static void empty_ray(uint32_t p, float* ox, float* oy, float* oz, float* tn, float* dx,
                      float* dy, float* dz, float* tx) {
    ox[p] = 0.0f; oy[p] = 0.0f; oz[p] = 0.0f; tn[p] = 1.0f;
    dx[p] = 0.0f; dy[p] = 0.0f; dz[p] = 1.0f; tx[p] = 0.0f;
}

void or_ao_rays(const float* pos, const uint32_t* idx,
                const float* pox, const float* poy, const float* poz,
                const float* pdx, const float* pdy, const float* pdz,
                const uint32_t* prim, const float* t, uint32_t w, uint32_t h,
                uint32_t frame, float offset, float tmin, float tmax,
                float* ox, float* oy, float* oz, float* tn,
                float* dx, float* dy, float* dz, float* tx) {
    for (uint32_t p = 0; p < w * h; ++p) {
        if (prim[p] == 0xFFFFFFFFu) {
            empty_ray(p, ox, oy, oz, tn, dx, dy, dz, tx);
            continue;
        }
        float n[3];
        face_normal(pos, idx, prim[p], pdx[p], pdy[p], pdz[p], n);
        ao_ray(p, n, pox[p] + t[p] * pdx[p], poy[p] + t[p] * pdy[p], poz[p] + t[p] * pdz[p], w, h,
               frame, offset, tmin, tmax, ox, oy, oz, tn, dx, dy, dz, tx);
    }
}

// AO rays over a two-level scene (the device's k_ao_scene): the normal of
// instance inst[p]'s hit triangle, oriented against the mapped direction,
// mapped back by M^T for a general instance (kind 0).
void or_ao_rays_scene(const or_instance* insts, uint32_t count,
                      const float* pox, const float* poy, const float* poz,
                      const float* pdx, const float* pdy, const float* pdz,
                      const uint32_t* prim, const uint32_t* inst, const float* t, uint32_t w,
                      uint32_t h, uint32_t frame, float offset, float tmin, float tmax,
                      float* ox, float* oy, float* oz, float* tn,
                      float* dx, float* dy, float* dz, float* tx) {
    for (uint32_t p = 0; p < w * h; ++p) {
        if (prim[p] == 0xFFFFFFFFu || inst[p] >= count) {
            empty_ray(p, ox, oy, oz, tn, dx, dy, dz, tx);
            continue;
        }
        const or_instance& I = insts[inst[p]];
        float n[3];
        if (I.kind == 0) {
            const float* m = I.world_to_object;
            const float ex = m[0] * pdx[p] + m[1] * pdy[p] + m[2] * pdz[p];
            const float ey = m[4] * pdx[p] + m[5] * pdy[p] + m[6] * pdz[p];
            const float ez = m[8] * pdx[p] + m[9] * pdy[p] + m[10] * pdz[p];
            float q[3];
            face_normal(I.pos, I.idx, prim[p], ex, ey, ez, q);
            n[0] = m[0] * q[0] + m[4] * q[1] + m[8] * q[2];
            n[1] = m[1] * q[0] + m[5] * q[1] + m[9] * q[2];
            n[2] = m[2] * q[0] + m[6] * q[1] + m[10] * q[2];
            normalize3(n[0], n[1], n[2]);
        } else {
            face_normal(I.pos, I.idx, prim[p], pdx[p], pdy[p], pdz[p], n);
        }
        ao_ray(p, n, pox[p] + t[p] * pdx[p], poy[p] + t[p] * pdy[p], poz[p] + t[p] * pdz[p], w, h,
               frame, offset, tmin, tmax, ox, oy, oz, tn, dx, dy, dz, tx);
    }
}

// Origins uniform in [lo,hi), directions uniform on the unit sphere
// (z = 1 - 2u, phi = 2 pi v), per-ray seeded hash streams.
void or_random_rays(uint32_t count, uint32_t seed, const float lo[3], const float hi[3],
                    float tmax, float* ox, float* oy, float* oz, float* tn,
                    float* dx, float* dy, float* dz, float* tx) {
    for (uint32_t i = 0; i < count; ++i) {
        float u0 = hash01(seed, i * 5 + 0), u1 = hash01(seed, i * 5 + 1), u2 = hash01(seed, i * 5 + 2);
        float u3 = hash01(seed, i * 5 + 3), u4 = hash01(seed, i * 5 + 4);
        ox[i] = lo[0] + u0 * (hi[0] - lo[0]);
        oy[i] = lo[1] + u1 * (hi[1] - lo[1]);
        oz[i] = lo[2] + u2 * (hi[2] - lo[2]);
        float z = 1.0f - 2.0f * u3;
        float r = std::sqrt(std::max(0.0f, 1.0f - z * z));
        float phi = 6.2831853f * u4;
        dx[i] = r * std::cos(phi); dy[i] = r * std::sin(phi); dz[i] = z;
        tn[i] = 0.0f; tx[i] = tmax;
    }
}

// ---------------------------------------------------------------------------
// LBVH: centroid bounds, 30-bit Morton codes, stable radix sort, implicit
// complete 4-ary tree in the fast_bvh_builder layout.
// ---------------------------------------------------------------------------

static inline void tri_vertices(const float* pos, const uint32_t* idx, uint32_t i,
                                const float** p0, const float** p1, const float** p2) {
    if (idx) {
        *p0 = pos + 3 * idx[i * 3 + 0];
        *p1 = pos + 3 * idx[i * 3 + 1];
        *p2 = pos + 3 * idx[i * 3 + 2];
    } else {  // unindexed: 9 floats per triangle (fast_bvh_builder.comp:120-122)
        *p0 = pos + 9 * (size_t)i;
        *p1 = *p0 + 3;
        *p2 = *p0 + 6;
    }
}

// Centroid (p0 + p1 + p2) / 3 as software_bvh.hpp:53; bounds = min/max.
void or_centroid_bounds(const float* pos, const uint32_t* idx, uint32_t n, float b[6]) {
    b[0] = b[1] = b[2] = FLT_MAX;
    b[3] = b[4] = b[5] = -FLT_MAX;
    for (uint32_t i = 0; i < n; ++i) {
        const float *p0, *p1, *p2;
        tri_vertices(pos, idx, i, &p0, &p1, &p2);
        for (int c = 0; c < 3; ++c) {
            float ce = (p0[c] + p1[c] + p2[c]) / 3.0f;
            b[c] = std::min(b[c], ce);
            b[3 + c] = std::max(b[3 + c], ce);
        }
    }
}

// Spread the low 10 bits of v to every third bit.
uint32_t or_expand_bits10(uint32_t v) {
    v &= 0x3FFu;
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// Morton-30: q_a = clamp(floor((c_a - lo_a) * s_a), 0, 1023) with
// s_a = 1024 / (hi_a - lo_a) (0 when the axis is flat);
// code = x bits at 3k+2, y at 3k+1, z at 3k.
void or_morton30(const float* pos, const uint32_t* idx, uint32_t n, const float b[6],
                 uint32_t* codes) {
    float s[3];
    for (int c = 0; c < 3; ++c) {
        float ext = b[3 + c] - b[c];
        s[c] = ext > 0.0f ? 1024.0f / ext : 0.0f;
    }
    for (uint32_t i = 0; i < n; ++i) {
        const float *p0, *p1, *p2;
        tri_vertices(pos, idx, i, &p0, &p1, &p2);
        uint32_t q[3];
        for (int c = 0; c < 3; ++c) {
            float ce = (p0[c] + p1[c] + p2[c]) / 3.0f;
            float f = (ce - b[c]) * s[c];
            f = std::min(std::max(f, 0.0f), 1023.0f);
            q[c] = (uint32_t)f;
        }
        codes[i] = (or_expand_bits10(q[0]) << 2) | (or_expand_bits10(q[1]) << 1) |
                   or_expand_bits10(q[2]);
    }
}

// Stable LSD radix sort of (key, value) pairs, 8-bit digits over 30 bits.
void or_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t n) {
    std::vector<uint32_t> k2(n), v2(n);
    uint32_t *ks = keys, *vs = vals, *kd = k2.data(), *vd = v2.data();
    for (int shift = 0; shift < 32; shift += 8) {
        uint32_t cnt[257] = {0};
        for (uint32_t i = 0; i < n; ++i) cnt[((ks[i] >> shift) & 0xFF) + 1]++;
        for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t p = cnt[(ks[i] >> shift) & 0xFF]++;
            kd[p] = ks[i];
            vd[p] = vs[i];
        }
        std::swap(ks, kd);
        std::swap(vs, vd);
    }
    // four passes: data is back in the caller's arrays
}

// include/radv_bvh.hpp:193-215 compute_tree_layout: leaf boxes ceil(n/4), then
// ceil(/4) until one box; stored root first; offsets are prefix sums;
// size = 192 + 128 boxes + 64 n.  (No 8-level cap: the arrays hold 32.)
void or_tree_layout(uint32_t n, uint32_t* counts, uint32_t* offsets, uint32_t* levels,
                    uint32_t* total_boxes, uint64_t* total_as) {
    uint32_t c[32];
    uint32_t L = 0;
    c[L++] = (n + 3) / 4;
    while (c[L - 1] > 1) { c[L] = (c[L - 1] + 3) / 4; L++; }
    for (uint32_t i = 0; i < L; ++i) counts[i] = c[L - 1 - i];
    uint32_t off = 0;
    for (uint32_t i = 0; i < L; ++i) { offsets[i] = off; off += counts[i]; }
    *levels = L;
    *total_boxes = off;
    *total_as = 192ull + 128ull * off + 64ull * n;
}

static void set_slot(or_bvh4_node& nd, int c, const float bb[6], int32_t child, int32_t count) {
    nd.minX[c] = bb[0]; nd.minY[c] = bb[1]; nd.minZ[c] = bb[2];
    nd.maxX[c] = bb[3]; nd.maxY[c] = bb[4]; nd.maxZ[c] = bb[5];
    nd.child[c] = child;
    nd.count[c] = count;
}

// Fill leaf records from sorted ids and compute every box bottom-up.
static void lbvh_fill(const float* pos, const uint32_t* idx, uint32_t n,
                      const uint32_t* sorted, or_bvh4_node* nodes, or_tri64* tris) {
    uint32_t counts[32], offs[32], L, tb;
    uint64_t sz;
    or_tree_layout(n, counts, offs, &L, &tb, &sz);
    std::vector<float> box((size_t)tb * 6);
    for (uint32_t i = 0; i < n; ++i) {
        const float *p0, *p1, *p2;
        tri_vertices(pos, idx, sorted[i], &p0, &p1, &p2);
        or_tri64& r = tris[i];
        std::memset(&r, 0, sizeof r);
        for (int c = 0; c < 3; ++c) { r.v[c] = p0[c]; r.v[3 + c] = p1[c]; r.v[6 + c] = p2[c]; }
        r.triangle_id = sorted[i];
        r.geometry_id_and_flags = 0x80000000u;
        r.id = 9u;
    }
    const uint32_t leafOff = offs[L - 1];
    for (uint32_t b = 0; b < counts[L - 1]; ++b) {  // leaf boxes: one triangle per slot
        or_bvh4_node& nd = nodes[leafOff + b];
        std::memset(&nd, 0, sizeof nd);
        float u[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (int c = 0; c < 4; ++c) {
            uint32_t s = 4 * b + c;
            if (s >= n) break;
            const float* v = tris[s].v;
            float bb[6];
            for (int a = 0; a < 3; ++a) {
                // + 0.0f canonicalises -0 to +0 (min/max of (+0, -0) may
                // return either zero; the device path adds the same 0).
                bb[a] = fmin3(v[a], v[3 + a], v[6 + a]) + 0.0f;
                bb[3 + a] = fmax3(v[a], v[3 + a], v[6 + a]) + 0.0f;
                u[a] = std::min(u[a], bb[a]);
                u[3 + a] = std::max(u[3 + a], bb[3 + a]);
            }
            set_slot(nd, c, bb, (int32_t)~s, 1);
        }
        std::memcpy(&box[(size_t)(leafOff + b) * 6], u, sizeof u);
    }
    for (int lv = (int)L - 2; lv >= 0; --lv) {  // fast_bvh_builder.comp:166-212 order
        for (uint32_t i = 0; i < counts[lv]; ++i) {
            or_bvh4_node& nd = nodes[offs[lv] + i];
            std::memset(&nd, 0, sizeof nd);
            float u[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
            for (int c = 0; c < 4; ++c) {
                uint32_t ch = 4 * i + c;
                if (ch >= counts[lv + 1]) break;
                uint32_t g = offs[lv + 1] + ch;
                const float* bb = &box[(size_t)g * 6];
                for (int a = 0; a < 3; ++a) {
                    u[a] = std::min(u[a], bb[a]);
                    u[3 + a] = std::max(u[3 + a], bb[3 + a]);
                }
                set_slot(nd, c, bb, (int32_t)g, -1);
            }
            std::memcpy(&box[(size_t)(offs[lv] + i) * 6], u, sizeof u);
        }
    }
}

// Full LBVH: bounds -> Morton -> stable sort by (code, index) -> implicit
// Box4 tree.  Node i of level l has children 4i..4i+3 of level l+1; leaf-level
// slots hold one triangle each (count 1, child = ~sorted position), inner
// slots count -1.  Empty slots are all-zero (as SoftwareBVH4 leaves them).
int or_lbvh_build(const float* pos, const uint32_t* idx, uint32_t n, or_bvh4_node* nodes,
                  uint32_t cap, uint32_t* node_count, or_tri64* tris, uint32_t* sorted_ids,
                  uint32_t* codes_sorted) {
    if (n == 0) { *node_count = 0; return 0; }
    uint32_t counts[32], offs[32], L, tb;
    uint64_t sz;
    or_tree_layout(n, counts, offs, &L, &tb, &sz);
    if (tb > cap) return 1;
    float b[6];
    or_centroid_bounds(pos, idx, n, b);
    std::vector<uint32_t> keys(n), vals(n);
    or_morton30(pos, idx, n, b, keys.data());
    for (uint32_t i = 0; i < n; ++i) vals[i] = i;
    or_sort_pairs(keys.data(), vals.data(), n);
    lbvh_fill(pos, idx, n, vals.data(), nodes, tris);
    if (sorted_ids) std::memcpy(sorted_ids, vals.data(), n * sizeof(uint32_t));
    if (codes_sorted) std::memcpy(codes_sorted, keys.data(), n * sizeof(uint32_t));
    *node_count = tb;
    return 0;
}

// ---------------------------------------------------------------------------
// Radix-tree LBVH: the binary radix tree over the Morton-sorted triangles
// (split of a sorted range where its highest differing key bit changes;
// equal keys split by the highest differing bit of the sorted position),
// ranges of <= max_prims triangles become leaves, and the tree is collapsed
// to Box4 with SoftwareBVH4::collapse (software_bvh.hpp:202-291) unchanged.
// ---------------------------------------------------------------------------

// Highest differing bit of positions i and j in the sorted sequence: key
// bits first, then 32 + the position bits (Karras 2012, delta with ties).
static inline int delta_key(const uint32_t* k, int64_t i, int64_t j) {
    uint32_t x = k[i] ^ k[j];
    if (x) return __builtin_clz(x);
    return 32 + __builtin_clz((uint32_t)(i ^ j));
}

// Split position of sorted range [f, l] (f < l): the last position whose
// prefix with f is longer than the prefix shared by the whole range.
static uint32_t find_split(const uint32_t* k, uint32_t f, uint32_t l) {
    const int common = delta_key(k, f, l);
    uint32_t split = f;
    uint32_t step = l - f;
    do {
        step = (step + 1) >> 1;
        uint32_t ns = split + step;
        if (ns < l && delta_key(k, f, ns) > common) split = ns;
    } while (step > 1);
    return split;
}

struct KarrasBuilder {
    const uint32_t* keys;
    uint32_t max_prims;
    std::vector<or_bvh2_node> nodes;
    const float* box;  // per sorted position: min xyz, max xyz

    int32_t build(uint32_t f, uint32_t l) {
        int32_t id = (int32_t)nodes.size();
        nodes.emplace_back();
        float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        if (l - f + 1 <= max_prims) {
            for (uint32_t i = f; i <= l; ++i)
                for (int a = 0; a < 3; ++a) {
                    mn[a] = std::min(mn[a], box[i * 6 + a]);
                    mx[a] = std::max(mx[a], box[i * 6 + 3 + a]);
                }
            nodes[id].leftChild = (int32_t)~f;
            nodes[id].rightChild = (int32_t)(l - f + 1);
        } else {
            uint32_t s = find_split(keys, f, l);
            int32_t a = build(f, s);
            int32_t b = build(s + 1, l);
            for (int c = 0; c < 3; ++c) {
                mn[c] = std::min(nodes[a].aabbMin[c], nodes[b].aabbMin[c]);
                mx[c] = std::max(nodes[a].aabbMax[c], nodes[b].aabbMax[c]);
            }
            nodes[id].leftChild = a;
            nodes[id].rightChild = b;
        }
        for (int c = 0; c < 3; ++c) {
            nodes[id].aabbMin[c] = mn[c];
            nodes[id].aabbMax[c] = mx[c];
        }
        return id;
    }
};

// Full radix-tree build: Morton codes, stable sort, radix tree with
// max_prims leaves (BVH2, pre-order), `rounds` rounds of treelet
// restructuring, collapse to Box4, 64-B leaves in sorted order.  Returns 1
// when a capacity is too small.
int or_rlbvh_build_opt(const float* pos, const uint32_t* idx, uint32_t n, uint32_t max_prims,
                       uint32_t rounds, or_bvh2_node* n2_out, uint32_t cap2, uint32_t* count2,
                       or_bvh4_node* n4_out, uint32_t cap4, uint32_t* count4, or_tri64* tris,
                       uint32_t* sorted_ids, uint32_t* codes_sorted) {
    *count2 = 0;
    *count4 = 0;
    if (n == 0) return 0;
    float b[6];
    or_centroid_bounds(pos, idx, n, b);
    std::vector<uint32_t> keys(n), vals(n);
    or_morton30(pos, idx, n, b, keys.data());
    for (uint32_t i = 0; i < n; ++i) vals[i] = i;
    or_sort_pairs(keys.data(), vals.data(), n);
    std::vector<float> box((size_t)n * 6);
    for (uint32_t i = 0; i < n; ++i) {
        const float *p0, *p1, *p2;
        tri_vertices(pos, idx, vals[i], &p0, &p1, &p2);
        or_tri64& r = tris[i];
        std::memset(&r, 0, sizeof r);
        for (int c = 0; c < 3; ++c) {
            r.v[c] = p0[c]; r.v[3 + c] = p1[c]; r.v[6 + c] = p2[c];
            box[i * 6 + c] = fmin3(p0[c], p1[c], p2[c]) + 0.0f;
            box[i * 6 + 3 + c] = fmax3(p0[c], p1[c], p2[c]) + 0.0f;
        }
        r.triangle_id = vals[i];
        r.geometry_id_and_flags = 0x80000000u;
        r.id = 9u;
    }
    KarrasBuilder kb{keys.data(), max_prims, {}, box.data()};
    kb.nodes.reserve(2 * (size_t)n);
    kb.build(0, n - 1);
    if (rounds) {
        TreeletOpt opt{kb.nodes, std::vector<float>(kb.nodes.size(), 0.0f)};
        for (uint32_t r = 0; r < rounds; ++r) opt.post(0);
    }
    if (kb.nodes.size() > cap2) return 1;
    std::memcpy(n2_out, kb.nodes.data(), kb.nodes.size() * sizeof(or_bvh2_node));
    *count2 = (uint32_t)kb.nodes.size();
    if (or_bvh4_collapse(n2_out, *count2, n4_out, cap4, count4)) return 1;
    if (sorted_ids) std::memcpy(sorted_ids, vals.data(), n * sizeof(uint32_t));
    if (codes_sorted) std::memcpy(codes_sorted, keys.data(), n * sizeof(uint32_t));
    return 0;
}

int or_rlbvh_build(const float* pos, const uint32_t* idx, uint32_t n, uint32_t max_prims,
                   or_bvh2_node* n2_out, uint32_t cap2, uint32_t* count2, or_bvh4_node* n4_out,
                   uint32_t cap4, uint32_t* count4, or_tri64* tris, uint32_t* sorted_ids,
                   uint32_t* codes_sorted) {
    return or_rlbvh_build_opt(pos, idx, n, max_prims, 0, n2_out, cap2, count2, n4_out, cap4,
                              count4, tris, sorted_ids, codes_sorted);
}

// Refit: same triangle order (taken from tris[i].triangle_id), new positions.
int or_lbvh_refit(const float* pos, const uint32_t* idx, uint32_t n, or_bvh4_node* nodes,
                  uint32_t node_count, or_tri64* tris) {
    (void)node_count;
    std::vector<uint32_t> ids(n);
    for (uint32_t i = 0; i < n; ++i) ids[i] = tris[i].triangle_id;
    lbvh_fill(pos, idx, n, ids.data(), nodes, tris);
    return 0;
}

// Refit of any Box4 tree whose inner children have larger indices than their
// parents (pre-order numbering of the radix / SoftwareBVH4 trees, level order
// of the implicit one): every leaf record re-read from its triangle_id, every
// slot box recomputed -- a leaf range as the union of its records' boxes
// (canonical +0 zeros, as the builds), an inner slot as the union of the
// child's occupied slots -- visiting nodes from the last to the first, so a
// child is final before its parent reads it.  Topology (child, count) and
// empty slots are unchanged.  Returns 1 if a child index does not exceed its
// parent's.  No reference counterpart: the reference rebuilds deforming
// meshes every frame (src/main.cpp:871-885).
int or_bvh4_refit(const float* pos, const uint32_t* idx, or_bvh4_node* nodes,
                  uint32_t node_count, or_tri64* tris, uint32_t ntris) {
    for (uint32_t s = 0; s < ntris; ++s) {
        const float *p0, *p1, *p2;
        tri_vertices(pos, idx, tris[s].triangle_id, &p0, &p1, &p2);
        for (int c = 0; c < 3; ++c) {
            tris[s].v[c] = p0[c];
            tris[s].v[3 + c] = p1[c];
            tris[s].v[6 + c] = p2[c];
        }
    }
    for (int64_t i = (int64_t)node_count - 1; i >= 0; --i) {
        or_bvh4_node& nd = nodes[i];
        for (int k = 0; k < 4; ++k) {
            float u[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
            if (nd.count[k] > 0) {
                const uint32_t first = (uint32_t)~nd.child[k];
                for (int32_t q = 0; q < nd.count[k]; ++q) {
                    const float* v = tris[first + q].v;
                    for (int a = 0; a < 3; ++a) {
                        u[a] = std::min(u[a], fmin3(v[a], v[3 + a], v[6 + a]) + 0.0f);
                        u[3 + a] = std::max(u[3 + a], fmax3(v[a], v[3 + a], v[6 + a]) + 0.0f);
                    }
                }
            } else if (nd.count[k] < 0) {
                if (nd.child[k] <= i || (uint32_t)nd.child[k] >= node_count) return 1;
                const or_bvh4_node& c = nodes[nd.child[k]];
                for (int q = 0; q < 4; ++q) {
                    if (c.count[q] == 0) continue;
                    u[0] = std::min(u[0], c.minX[q]); u[3] = std::max(u[3], c.maxX[q]);
                    u[1] = std::min(u[1], c.minY[q]); u[4] = std::max(u[4], c.maxY[q]);
                    u[2] = std::min(u[2], c.minZ[q]); u[5] = std::max(u[5], c.maxZ[q]);
                }
            } else {
                continue;
            }
            set_slot(nd, k, u, nd.child[k], nd.count[k]);
        }
    }
    return 0;
}

// shaders/fast_bvh_builder.comp:91-257 byte image: 136-B header (+ pre-root
// links at 0xB4/0xBC) at 0, Box32 nodes at 192 + 128 i (children
// (rel >> 3) | type, rel measured from offset 192), 64-B triangle nodes after
// the boxes.  order[j] (NULL = identity, the reference behaviour) selects which
// input triangle lands in leaf j; triangle_id = that input index.
int or_radv_encode(const float* xyz9, uint32_t n, const uint32_t* order, uint8_t* out,
                   uint64_t out_bytes) {
    uint32_t counts[32], offs[32], L, tb;
    uint64_t sz;
    or_tree_layout(n, counts, offs, &L, &tb, &sz);
    if (out_bytes < sz) return 1;
    std::memset(out, 0, sz);
    auto put32 = [&](uint64_t off, uint32_t v) { std::memcpy(out + off, &v, 4); };
    auto putf = [&](uint64_t off, float v) { std::memcpy(out + off, &v, 4); };
    auto put64 = [&](uint64_t off, uint64_t v) { std::memcpy(out + off, &v, 8); };
    const float qnan = []() { uint32_t b = 0x7fc00000u; float f; std::memcpy(&f, &b, 4); return f; }();
    const uint64_t boxBase = 192, triBase = 192 + 128ull * tb;
    std::vector<float> bmin((size_t)tb * 3), bmax((size_t)tb * 3);
    const uint32_t leafOff = offs[L - 1];
    for (uint32_t t = 0; t < counts[L - 1]; ++t) {  // phase 1 :106-161
        uint64_t nb = boxBase + 128ull * (leafOff + t);
        float um[3] = {1e30f, 1e30f, 1e30f}, uM[3] = {-1e30f, -1e30f, -1e30f};
        for (uint32_t c = 0; c < 4; ++c) {
            uint32_t j = 4 * t + c;
            uint64_t cb = nb + 16 + 24ull * c;
            if (j < n) {
                uint32_t src = order ? order[j] : j;
                const float* v = xyz9 + 9ull * src;
                uint64_t tn = triBase + 64ull * j;
                for (int k = 0; k < 9; ++k) putf(tn + 4 * k, v[k]);
                put32(tn + 48, src);
                put32(tn + 52, 0x80000000u);
                put32(tn + 60, 9u);
                put32(nb + 4 * c, (uint32_t)((128ull * tb + 64ull * j) >> 3) | 0u);
                for (int a = 0; a < 3; ++a) {
                    // zeros canonicalised to +0, as in lbvh_fill
                    float mn = std::min(v[a], std::min(v[3 + a], v[6 + a])) + 0.0f;
                    float mx = std::max(v[a], std::max(v[3 + a], v[6 + a])) + 0.0f;
                    putf(cb + 4 * a, mn);
                    putf(cb + 12 + 4 * a, mx);
                    um[a] = std::min(um[a], mn);
                    uM[a] = std::max(uM[a], mx);
                }
            } else {
                put32(nb + 4 * c, 0xFFFFFFFFu);
                for (int a = 0; a < 6; ++a) putf(cb + 4 * a, qnan);
            }
        }
        for (int a = 0; a < 3; ++a) {
            bmin[(leafOff + t) * 3 + a] = um[a];
            bmax[(leafOff + t) * 3 + a] = uM[a];
        }
    }
    for (int lv = (int)L - 2; lv >= 0; --lv) {  // phase 2 :166-212
        for (uint32_t t = 0; t < counts[lv]; ++t) {
            uint32_t p = offs[lv] + t;
            uint64_t nb = boxBase + 128ull * p;
            float um[3] = {1e30f, 1e30f, 1e30f}, uM[3] = {-1e30f, -1e30f, -1e30f};
            for (uint32_t c = 0; c < 4; ++c) {
                uint32_t cl = 4 * t + c;
                uint64_t cb = nb + 16 + 24ull * c;
                if (cl < counts[lv + 1]) {
                    uint32_t g = offs[lv + 1] + cl;
                    put32(nb + 4 * c, ((g * 128u) >> 3) | 5u);
                    for (int a = 0; a < 3; ++a) {
                        putf(cb + 4 * a, bmin[g * 3 + a]);
                        putf(cb + 12 + 4 * a, bmax[g * 3 + a]);
                        um[a] = std::min(um[a], bmin[g * 3 + a]);
                        uM[a] = std::max(uM[a], bmax[g * 3 + a]);
                    }
                } else {
                    put32(nb + 4 * c, 0xFFFFFFFFu);
                    for (int a = 0; a < 6; ++a) putf(cb + 4 * a, qnan);
                }
            }
            for (int a = 0; a < 3; ++a) { bmin[p * 3 + a] = um[a]; bmax[p * 3 + a] = uM[a]; }
        }
    }
    // phase 3 :220-256 header + pre-root links
    put32(0x00, 192u);
    put32(0x04, 0u);
    for (int a = 0; a < 3; ++a) { putf(0x08 + 4 * a, bmin[a]); putf(0x14 + 4 * a, bmax[a]); }
    put32(0x20, 8u); put32(0x24, 1u); put32(0x28, 1u);
    put64(0x30, sz);
    put64(0x38, sz + 128);
    put32(0x40, 1u); put32(0x44, 1u); put32(0x48, 1u);
    put64(0x50, sz);
    put32(0x58, 0u);
    put32(0x5C, 1u);
    put64(0x60, 192ull + 128ull * tb);
    put64(0x68, n);
    put64(0x70, 0u);
    put32(0x78, 0u);
    put32(0x7C, 4u);
    put32(0x80, 1u);
    put32(180, 5u);
    put32(188, 0xFFFFFFFFu);
    return 0;
}

// Decoder of a RADV acceleration-structure image (include/radv_bvh.hpp:22-152,
// the header at 0, Box32 nodes from bvh_offset, triangle nodes from
// leaf_nodes_offset) into Box4 arrays: Box32 node i (at bvh_offset + 128 i)
// becomes Box4 node i; a child id (offset >> 3) | type (offset from
// bvh_offset, radv_bvh.hpp:39-49) of type box32 becomes an inner slot (count
// -1, child = offset / 128), of type triangle a one-triangle leaf slot
// (count 1, child = ~j for the triangle node j at leaf_nodes_offset + 64 j),
// 0xFFFFFFFF an all-zero empty slot; slot boxes are copied.  Triangle nodes
// are copied whole (the 64-B leaf record).  Returns 1 on a malformed image
// (sizes, alignment, a child outside its section or of another type).
int or_radv_decode(const uint8_t* img, uint64_t bytes, or_bvh4_node* nodes, uint32_t cap,
                   uint32_t* node_count, or_tri64* tris, uint32_t tcap, uint32_t* tri_count) {
    if (bytes < 136) return 1;
    uint32_t bvh;
    uint64_t leaf, prims;
    std::memcpy(&bvh, img + 0x00, 4);
    std::memcpy(&leaf, img + 0x60, 8);
    std::memcpy(&prims, img + 0x68, 8);
    if (bvh % 64 || leaf < bvh || (leaf - bvh) % 128 || leaf + 64 * prims > bytes) return 1;
    const uint64_t boxes = (leaf - bvh) / 128;
    if (boxes > cap || prims > tcap) return 1;
    for (uint64_t i = 0; i < boxes; ++i) {
        const uint8_t* b = img + bvh + 128 * i;
        or_bvh4_node& nd = nodes[i];
        std::memset(&nd, 0, sizeof nd);
        for (int c = 0; c < 4; ++c) {
            uint32_t id;
            float bb[6];
            std::memcpy(&id, b + 4 * c, 4);
            std::memcpy(bb, b + 16 + 24 * c, 24);
            if (id == 0xFFFFFFFFu) continue;
            const uint64_t off = (uint64_t)(id & ~7u) << 3;
            const uint32_t type = id & 7u;
            int32_t child, count;
            if (type == 5) {
                if (off % 128 || off / 128 >= boxes) return 1;
                child = (int32_t)(off / 128);
                count = -1;
            } else if (type == 0) {
                const uint64_t tbase = leaf - bvh;
                if (off < tbase || (off - tbase) % 64 || (off - tbase) / 64 >= prims) return 1;
                child = (int32_t)~(uint32_t)((off - tbase) / 64);
                count = 1;
            } else {
                return 1;
            }
            set_slot(nd, c, bb, child, count);  // AABB = min xyz, max xyz
        }
    }
    std::memcpy(tris, img + leaf, 64 * prims);
    *node_count = (uint32_t)boxes;
    *tri_count = (uint32_t)prims;
    return 0;
}

// ---------------------------------------------------------------------------
// Post-processing and metrics
// ---------------------------------------------------------------------------

// 1 SPP AO (rq_ao.comp:91-96): 1 where the AO ray is unoccluded, 0 where
// occluded; pixels whose primary ray missed are 1 (rq_ao.comp:53-55).
void or_ao_resolve(const uint32_t* bits, const uint32_t* prim, uint32_t npix, float* ao) {
    for (uint32_t p = 0; p < npix; ++p) {
        bool occ = (bits[p >> 5] >> (p & 31)) & 1u;
        ao[p] = (prim[p] == 0xFFFFFFFFu) ? 1.0f : (occ ? 0.0f : 1.0f);
    }
}

// 7x7 depth-gated bilateral filter.  Spatial weights are the products
// w[|i|] w[|j|] of the first four taps of bilateral_blur.comp:28-31; a
// neighbour counts iff |depth - centre depth| < thr (bilateral_blur.comp:41,49)
// and lies inside the image; result = sum / max(total, 0.001).  Pixels with
// depth >= 1e30 (background) pass through (bilateral_blur.comp:23-26).
void or_blur7(const float* in, const float* depth, uint32_t w, uint32_t h, float thr,
              float* out) {
    const float wt[4] = {0.227027f, 0.1945946f, 0.1216216f, 0.054054f};
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            uint32_t p = y * w + x;
            float cd = depth[p];
            if (cd >= 1e30f) { out[p] = in[p]; continue; }
            float acc = 0.0f, tot = 0.0f;
            for (int j = -3; j <= 3; ++j) {
                int yy = (int)y + j;
                if (yy < 0 || yy >= (int)h) continue;
                for (int i = -3; i <= 3; ++i) {
                    int xx = (int)x + i;
                    if (xx < 0 || xx >= (int)w) continue;
                    uint32_t q = (uint32_t)yy * w + (uint32_t)xx;
                    if (!(std::fabs(depth[q] - cd) < thr)) continue;
                    float ww = wt[i < 0 ? -i : i] * wt[j < 0 ? -j : j];
                    acc = acc + in[q] * ww;
                    tot = tot + ww;
                }
            }
            out[p] = acc / std::max(tot, 0.001f);
        }
}

double or_rms(const float* a, const float* b, uint64_t n) {
    double s = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        double d = (double)a[i] - (double)b[i];
        s += d * d;
    }
    return n ? std::sqrt(s / (double)n) : 0.0;
}

// common/vk_timer.cpp:105-108: sorted[floor(q * (n - 1))] (q as float).
float or_percentile(float* v, uint32_t n, float q) {
    if (n == 0) return 0.0f;
    std::sort(v, v + n);
    size_t i = (size_t)(q * (float)(n - 1));
    return v[i];
}

// GpuTimer::runBenchmark statistics (common/vk_timer.cpp:96-121) of n
// samples (sorted in place).  out: median = sorted[n / 2] (:103), min, max,
// p5 and p95 = sorted[floor(q (n - 1))] with q as float (:105-108), mean =
// float(double sum / n) (:110-111), standard deviation over n with the float
// mean (:113-118), then p99 by the rule of p5 / p95.  n = 0: all zero.
void or_timer_stats(float* v, uint32_t n, float out[8]) {
    for (int i = 0; i < 8; ++i) out[i] = 0.0f;
    if (n == 0) return;
    std::sort(v, v + n);
    const auto pq = [&](float q) { return v[(size_t)(q * (float)(n - 1))]; };
    out[0] = v[n / 2];
    out[1] = v[0];
    out[2] = v[n - 1];
    out[3] = pq(0.05f);
    out[4] = pq(0.95f);
    double sum = 0.0;
    for (uint32_t i = 0; i < n; ++i) sum += v[i];
    const float avg = (float)(sum / n);
    out[5] = avg;
    double var = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        const double d = v[i] - avg;
        var += d * d;
    }
    out[6] = (float)std::sqrt(var / n);
    out[7] = pq(0.99f);
}

// ImageMetrics::compute (common/image_metrics.hpp:16-104) of two w x h
// float images.  out: mse = sum of (double(a) - double(b))^2 in pixel order
// / pixel count; psnr = 100 when mse < 1e-12, else 10 log10(1 / mse);
// ssim = mean over the full 8 x 8 blocks (row-major block order) of the
// block SSIM clamped to [0, 1]: block means and (n - 1)-normalised variances
// and covariance in double, each summed over the block's pixels in row-major
// order, C1 = 0.01^2, C2 = 0.03^2; 1.0 when no full block fits.  No pixels:
// mse 0, psnr 100, ssim 1.  block (optional): each block's clamped SSIM.
void or_image_metrics(const float* a, const float* b, uint32_t w, uint32_t h, double out[3],
                      double* block) {
    out[0] = 0.0;
    out[1] = 100.0;
    out[2] = 1.0;
    const size_t npix = (size_t)w * h;
    if (npix == 0) return;
    double sq = 0.0;
    for (size_t i = 0; i < npix; ++i) {
        const double d = (double)a[i] - (double)b[i];
        sq += d * d;
    }
    out[0] = sq / (double)npix;
    out[1] = out[0] < 1e-12 ? 100.0 : 10.0 * std::log10(1.0 / out[0]);
    const uint32_t bw = w / 8, bh = h / 8;
    if (bw == 0 || bh == 0) return;
    const double c1 = (0.01 * 1.0) * (0.01 * 1.0), c2 = (0.03 * 1.0) * (0.03 * 1.0);
    double total = 0.0;
    for (uint32_t by = 0; by < bh; ++by)
        for (uint32_t bx = 0; bx < bw; ++bx) {
            const float* pa = a + (size_t)by * 8 * w + bx * 8;
            const float* pb = b + (size_t)by * 8 * w + bx * 8;
            double ma = 0.0, mb = 0.0;
            for (int y = 0; y < 8; ++y)
                for (int x = 0; x < 8; ++x) {
                    ma += pa[(size_t)y * w + x];
                    mb += pb[(size_t)y * w + x];
                }
            ma /= 64;
            mb /= 64;
            double va = 0.0, vb = 0.0, cab = 0.0;
            for (int y = 0; y < 8; ++y)
                for (int x = 0; x < 8; ++x) {
                    const double da = pa[(size_t)y * w + x] - ma;
                    const double db = pb[(size_t)y * w + x] - mb;
                    va += da * da;
                    vb += db * db;
                    cab += da * db;
                }
            va /= 63;
            vb /= 63;
            cab /= 63;
            const double num = (2.0 * ma * mb + c1) * (2.0 * cab + c2);
            const double den = (ma * ma + mb * mb + c1) * (va + vb + c2);
            const double s = std::min(std::max(num / den, 0.0), 1.0);
            if (block) block[(size_t)by * bw + bx] = s;
            total += s;
        }
    out[2] = total / (double)((size_t)bw * bh);
}

}  // extern "C"
