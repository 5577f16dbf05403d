// bvh_gpu.hip - device BVH construction for one-level scenes (triangles, spheres, quads, lights).
//
// The reference builds its tree on one GPU thread with a random split axis and an O(n^2) sort
// (core/src/misc/bvh_node.cuh:63-108, from load_scene<<<1,1>>>, core/src/gpu.cu:481-751). Here the
// whole device places the primitives:
//   1. box and centroid of every primitive (source order), centroid bounds by block reductions;
//   2. 63-bit Morton codes of the centroids, radix sort of (code, primitive);
//   3. the binary radix tree over the sorted codes (Karras, "Maximizing parallelism in the
//      construction of BVHs, octrees, and k-d trees", HPG 2012): one thread per internal node,
//      equal codes ordered by position;
//   4. boxes fitted bottom-up, one thread per leaf climbing while it is the second child to arrive;
//   5. per-type ranks of the sorted primitives (exclusive scan): a leaf's records are a contiguous
//      run of its type's array in sorted order, so a leaf reference is (type, first rank, count);
//   6. a top-down pass per level turns the binary tree into the traversal's nodes: four-wide nodes
//      take up to four descendants by opening the largest-area child first (binary nodes take the
//      two children); new interior children get the next node slots and form the next level;
//   7. every primitive's records are written at its rank in the scene's geometry allocation.
// The tree differs from the host's binned-SAH tree (it is built for speed of construction, for
// scenes that arrive with a request); closest hits do not depend on the tree.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <thread>

#include "crt_kernels.h"
#include "crt_layout.h"
#include "crt_runtime.h"
#include "scene_build.h"

#define CRT_EMPTY_CHILD 0x7fffffff

namespace {

constexpr int kMaxLevels = 512;

struct TypeCount {
    uint32_t c[3];   // CRT_PRIM_TRI, CRT_PRIM_SPHERE, CRT_PRIM_QUAD
};
struct TypeCountAdd {
    __host__ __device__ TypeCount operator()(const TypeCount& a, const TypeCount& b) const {
        return TypeCount{{a.c[0] + b.c[0], a.c[1] + b.c[1], a.c[2] + b.c[2]}};
    }
};

// Source-order primitive index space: quads and lights [0, Q), spheres [Q, Q + S), triangles
// [Q + S, N).
struct Prims {
    const float4* quads;     // 5 per quad
    const float4* qboxes;    // 2 per quad record: lo, hi (computed on the host, shapes included)
    const float4* spheres;   // 2 per sphere
    const float4* tris;      // 3 per triangle
    uint32_t nq, ns, nt;
    __device__ __forceinline__ uint32_t type_of(uint32_t p) const {
        return p < nq ? CRT_PRIM_QUAD : (p < nq + ns ? CRT_PRIM_SPHERE : CRT_PRIM_TRI);
    }
};

__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// Box of one primitive, as the host build computes it (scene_build.cpp; quad and shape boxes come
// from the host).
__device__ __forceinline__ void prim_box(const Prims& P, uint32_t p, float3& lo, float3& hi) {
#pragma clang fp contract(off)
    if (p < P.nq) {
        const float4 a = P.qboxes[2 * (size_t)p], b = P.qboxes[2 * (size_t)p + 1];
        lo = make_float3(a.x, a.y, a.z);
        hi = make_float3(b.x, b.y, b.z);
    } else if (p < P.nq + P.ns) {
        const float4 s = P.spheres[2 * (size_t)(p - P.nq)];
        const float r = fabsf(s.w);
        lo = make_float3(s.x - r, s.y - r, s.z - r);
        hi = make_float3(s.x + r, s.y + r, s.z + r);
    } else {
        const float4* t = P.tris + 3 * (size_t)(p - P.nq - P.ns);
        const float4 a = t[0], b = t[1], c = t[2];
        lo = make_float3(fminf(a.x, fminf(b.x, c.x)), fminf(a.y, fminf(b.y, c.y)), fminf(a.z, fminf(b.z, c.z)));
        hi = make_float3(fmaxf(a.x, fmaxf(b.x, c.x)), fmaxf(a.y, fmaxf(b.y, c.y)), fmaxf(a.z, fmaxf(b.z, c.z)));
    }
}

// 1. primitive boxes; centroid bounds as order-preserving integers (cb[0..2] min, cb[3..5] max)
__global__ __launch_bounds__(256) void k_bvh_bounds(Prims P, uint32_t n, float4* __restrict__ plo, float4* __restrict__ phi,
                                                    uint32_t* __restrict__ cb) {
    __shared__ uint32_t s[6][4];
    float cmin[3] = {3.4e38f, 3.4e38f, 3.4e38f}, cmax[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    for (uint32_t p = blockIdx.x * 256u + threadIdx.x; p < n; p += gridDim.x * 256u) {
        float3 lo, hi;
        prim_box(P, p, lo, hi);
        plo[p] = make_float4(lo.x, lo.y, lo.z, __uint_as_float(P.type_of(p)));
        phi[p] = make_float4(hi.x, hi.y, hi.z, 0.0f);
        const float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
        for (int a = 0; a < 3; ++a) { cmin[a] = fminf(cmin[a], c[a]); cmax[a] = fmaxf(cmax[a], c[a]); }
    }
    for (int a = 0; a < 3; ++a) {
        for (int off = 32; off > 0; off >>= 1) {
            cmin[a] = fminf(cmin[a], __shfl_xor(cmin[a], off));
            cmax[a] = fmaxf(cmax[a], __shfl_xor(cmax[a], off));
        }
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int a = 0; a < 3; ++a) { s[a][wave] = f2ord(cmin[a]); s[3 + a][wave] = f2ord(cmax[a]); }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int a = 0; a < 3; ++a) {
            uint32_t mn = s[a][0], mx = s[3 + a][0];
            for (int w = 1; w < 4; ++w) { mn = min(mn, s[a][w]); mx = max(mx, s[3 + a][w]); }
            atomicMin(cb + a, mn);
            atomicMax(cb + 3 + a, mx);
        }
    }
}

__device__ __forceinline__ uint64_t spread21(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// 2. Morton code of each centroid in the centroid bounds (21 bits per axis)
__global__ __launch_bounds__(256) void k_bvh_morton(const float4* __restrict__ plo, const float4* __restrict__ phi,
                                                    const uint32_t* __restrict__ cb, uint32_t n, uint64_t* __restrict__ keys,
                                                    uint32_t* __restrict__ vals) {
    float lo[3], sc[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = ord2f(cb[a]);
        const float ext = ord2f(cb[3 + a]) - lo[a];
        sc[a] = ext > 0.0f ? 2097152.0f / ext : 0.0f;
    }
    for (uint32_t p = blockIdx.x * 256u + threadIdx.x; p < n; p += gridDim.x * 256u) {
        const float4 a = plo[p], b = phi[p];
        const float c[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
        uint32_t q[3];
        for (int k = 0; k < 3; ++k) {
            float t = (c[k] - lo[k]) * sc[k];
            t = fminf(fmaxf(t, 0.0f), 2097151.0f);
            q[k] = (uint32_t)t;
        }
        keys[p] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
        vals[p] = p;
    }
}

// common prefix length of sorted keys i and j (positions break ties between equal codes)
__device__ __forceinline__ int delta(const uint64_t* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint64_t a = k[i], b = k[j];
    if (a == b) return 64 + __clz((uint32_t)(i ^ j));
    return __clzll((long long)(a ^ b));
}

// 3. binary radix tree: internal node i covers sorted positions [first, first + count); children
// >= 0 are internal nodes, < 0 leaves (~position)
__global__ __launch_bounds__(256) void k_bvh_karras(const uint64_t* __restrict__ k, int n, int32_t* __restrict__ cl,
                                                    int32_t* __restrict__ cr, int32_t* __restrict__ parent,
                                                    int32_t* __restrict__ leaf_parent, uint32_t* __restrict__ first,
                                                    uint32_t* __restrict__ count) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n - 1; i += gridDim.x * 256) {
        const int d = (delta(k, n, i, i + 1) - delta(k, n, i, i - 1)) >= 0 ? 1 : -1;
        const int dmin = delta(k, n, i, i - d);
        int lmax = 2;
        while (delta(k, n, i, i + lmax * d) > dmin) lmax <<= 1;
        int l = 0;
        for (int t = lmax >> 1; t >= 1; t >>= 1)
            if (delta(k, n, i, i + (l + t) * d) > dmin) l += t;
        const int j = i + l * d;
        const int dnode = delta(k, n, i, j);
        int s = 0, t = l;
        do {
            t = (t + 1) >> 1;
            if (delta(k, n, i, i + (s + t) * d) > dnode) s += t;
        } while (t > 1);
        const int g = i + s * d + min(d, 0);
        const int lo = min(i, j), hi = max(i, j);
        const int32_t left = lo == g ? ~g : g, right = hi == g + 1 ? ~(g + 1) : g + 1;
        cl[i] = left;
        cr[i] = right;
        if (left < 0) leaf_parent[~left] = i; else parent[left] = i;
        if (right < 0) leaf_parent[~right] = i; else parent[right] = i;
        first[i] = (uint32_t)lo;
        count[i] = (uint32_t)(hi - lo + 1);
        if (i == 0) parent[0] = -1;
    }
}

// 4. boxes bottom-up: each leaf climbs; at every node the first of the two children to arrive stops,
// the second fits the node's box (and its type mask) from both children and goes on
__global__ __launch_bounds__(256) void k_bvh_refit(int n, const uint32_t* __restrict__ vals, const float4* __restrict__ plo,
                                                   const float4* __restrict__ phi, const int32_t* __restrict__ cl,
                                                   const int32_t* __restrict__ cr, const int32_t* __restrict__ parent,
                                                   const int32_t* __restrict__ leaf_parent, float4* nlo, float4* nhi,
                                                   uint32_t* tmask, uint32_t* __restrict__ visits) {
    for (int s = blockIdx.x * 256 + threadIdx.x; s < n; s += gridDim.x * 256) {
        int node = leaf_parent[s];
        while (node >= 0) {
            __threadfence();
            if (atomicAdd(visits + node, 1u) == 0u) break;
            __threadfence();
            float4 lo[2], hi[2];
            uint32_t tm = 0;
            const int32_t ch[2] = {cl[node], cr[node]};
            for (int c = 0; c < 2; ++c) {
                if (ch[c] < 0) {
                    const uint32_t p = vals[~ch[c]];
                    lo[c] = plo[p]; hi[c] = phi[p];
                    tm |= 1u << __float_as_uint(lo[c].w);
                } else {
                    // written by the thread that completed the child, before its fence and atomic
                    lo[c] = nlo[ch[c]];
                    hi[c] = nhi[ch[c]];
                    tm |= tmask[ch[c]];
                }
            }
            nlo[node] = make_float4(fminf(lo[0].x, lo[1].x), fminf(lo[0].y, lo[1].y), fminf(lo[0].z, lo[1].z), 0.0f);
            nhi[node] = make_float4(fmaxf(hi[0].x, hi[1].x), fmaxf(hi[0].y, hi[1].y), fmaxf(hi[0].z, hi[1].z), 0.0f);
            tmask[node] = tm;
            node = parent[node];
        }
    }
}

__global__ __launch_bounds__(256) void k_bvh_typeflags(const uint32_t* __restrict__ vals, Prims P, uint32_t n,
                                                       TypeCount* __restrict__ out) {
    for (uint32_t s = blockIdx.x * 256u + threadIdx.x; s < n; s += gridDim.x * 256u) {
        const uint32_t t = P.type_of(vals[s]);
        out[s] = TypeCount{{t == CRT_PRIM_TRI ? 1u : 0u, t == CRT_PRIM_SPHERE ? 1u : 0u, t == CRT_PRIM_QUAD ? 1u : 0u}};
    }
}

struct Tree {
    const int32_t* cl;
    const int32_t* cr;
    const float4* nlo;
    const float4* nhi;
    const uint32_t* first;
    const uint32_t* count;
    const uint32_t* tmask;
    const uint32_t* vals;
    const float4* plo;
    const float4* phi;
    const TypeCount* ranks;   // exclusive per-type ranks of the sorted positions
    Prims P;
    uint32_t leaf_size;
    __device__ __forceinline__ void box(int32_t c, float4& lo, float4& hi) const {
        if (c < 0) { const uint32_t p = vals[~c]; lo = plo[p]; hi = phi[p]; }
        else { lo = nlo[c]; hi = nhi[c]; }
    }
    __device__ __forceinline__ float area(int32_t c) const {
        float4 lo, hi;
        box(c, lo, hi);
        const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
    // becomes a leaf reference: a primitive, or a subtree of at most leaf_size primitives of one type
    __device__ __forceinline__ bool leafy(int32_t c) const {
        return c < 0 || (leaf_size > 1u && count[c] <= leaf_size && __popc(tmask[c]) == 1);
    }
    __device__ __forceinline__ int32_t leaf_ref(int32_t c, uint32_t& cnt, uint32_t* too_large) const {
        uint32_t f, type;
        if (c < 0) { f = (uint32_t)~c; cnt = 1u; type = P.type_of(vals[f]); }
        else { f = first[c]; cnt = count[c]; type = (uint32_t)__ffs(tmask[c]) - 1u; }
        const uint32_t r = ranks[f].c[type];
        if (r + cnt > CRT_LEAF_MAX_FIRST) { atomicOr(too_large, 1u); return CRT_LEAF_REF(type, 0u, 1u); }
        return CRT_LEAF_REF(type, r, cnt);
    }
};

// wave-wide exclusive prefix sum of v (all lanes active); returns the wave total in *total
__device__ __forceinline__ uint32_t wave_exclusive(uint32_t v, uint32_t& total) {
    const int lane = threadIdx.x & 63;
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, off);
        if (lane >= off) x += y;
    }
    total = (uint32_t)__shfl((int)x, 63);
    return x - v;
}

// 6. one level of the top-down collapse: output nodes [lb, le) (binary node of each in slot_b)
// take their children; new interior children get the slots behind *counter (the next level)
template <int WIDTH>
__global__ __launch_bounds__(256) void k_bvh_collapse(Tree T, const uint32_t* __restrict__ level_base, int level,
                                                      int32_t* __restrict__ slot_b, uint32_t* __restrict__ counter,
                                                      float4* __restrict__ out, double* __restrict__ sah, float inv_root_area,
                                                      uint32_t* __restrict__ too_large) {
    const uint32_t lb = level_base[level], le = level_base[level + 1];
    const uint32_t n = le - lb;
    const uint32_t iters = (n + gridDim.x * 256u - 1u) / (gridDim.x * 256u);
    for (uint32_t it = 0; it < iters; ++it) {
        const uint32_t i = it * gridDim.x * 256u + blockIdx.x * 256u + threadIdx.x;
        const bool act = i < n;
        int32_t ch[4] = {0, 0, 0, 0};
        int nc = 0;
        uint32_t nnew = 0, nleaf = 0;
        double cost = 0.0;
        if (act) {
            const int32_t b = slot_b[lb + i];
            ch[0] = T.cl[b]; ch[1] = T.cr[b];
            nc = 2;
            if (WIDTH == 4) {
                while (nc < 4) {
                    int best = -1;
                    float ba = -1.0f;
                    for (int k = 0; k < nc; ++k) {
                        if (T.leafy(ch[k])) continue;
                        const float a = T.area(ch[k]);
                        if (a > ba) { ba = a; best = k; }
                    }
                    if (best < 0) break;
                    const int32_t x = ch[best];
                    ch[best] = T.cl[x];
                    ch[nc++] = T.cr[x];
                }
            }
            for (int k = 0; k < nc; ++k) nnew += T.leafy(ch[k]) ? 0u : 1u;
            nleaf = (uint32_t)nc - nnew;
            cost = (double)T.area(b) * inv_root_area;
        }
        uint32_t total;
        const uint32_t pre = wave_exclusive(nnew, total);
        uint32_t base = 0;
        if ((threadIdx.x & 63) == 0 && total) base = atomicAdd(counter, total);
        base = (uint32_t)__shfl((int)base, 0) + pre;
        if (act) {
            const uint32_t slot = lb + i;
            float4 lo[4], hi[4];
            int32_t ref[4];
            for (int k = 0; k < 4; ++k) {
                const float inf = __builtin_inff();
                lo[k] = make_float4(inf, inf, inf, 0.0f);
                hi[k] = lo[k];
                ref[k] = CRT_EMPTY_CHILD;
            }
            for (int k = 0; k < nc; ++k) {
                T.box(ch[k], lo[k], hi[k]);
                if (T.leafy(ch[k])) {
                    uint32_t cnt;
                    ref[k] = T.leaf_ref(ch[k], cnt, too_large);
                    cost += (double)T.area(ch[k]) * inv_root_area * (double)cnt;
                } else {
                    slot_b[base] = ch[k];
                    ref[k] = (int32_t)base;
                    ++base;
                }
            }
            if (WIDTH == 4) {
                float4* o = out + 8 * (size_t)slot;
                o[0] = make_float4(lo[0].x, lo[1].x, lo[2].x, lo[3].x);
                o[1] = make_float4(hi[0].x, hi[1].x, hi[2].x, hi[3].x);
                o[2] = make_float4(lo[0].y, lo[1].y, lo[2].y, lo[3].y);
                o[3] = make_float4(hi[0].y, hi[1].y, hi[2].y, hi[3].y);
                o[4] = make_float4(lo[0].z, lo[1].z, lo[2].z, lo[3].z);
                o[5] = make_float4(hi[0].z, hi[1].z, hi[2].z, hi[3].z);
                o[6] = make_float4(__int_as_float(ref[0]), __int_as_float(ref[1]), __int_as_float(ref[2]), __int_as_float(ref[3]));
                o[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            } else {
                float4* o = out + 4 * (size_t)slot;
                o[0] = make_float4(lo[0].x, hi[0].x, lo[0].y, hi[0].y);
                o[1] = make_float4(lo[1].x, hi[1].x, lo[1].y, hi[1].y);
                o[2] = make_float4(lo[0].z, hi[0].z, lo[1].z, hi[1].z);
                o[3] = make_float4(__int_as_float(ref[0]), __int_as_float(ref[1]), 0.0f, 0.0f);
            }
        }
        // SAH cost of the emitted nodes and leaves, leaf count: one atomic per wave
        for (int off = 32; off > 0; off >>= 1) {
            cost += __shfl_xor(cost, off);
            nleaf += (uint32_t)__shfl_xor((int)nleaf, off);
        }
        if ((threadIdx.x & 63) == 0 && cost != 0.0) atomicAdd(sah, cost);
        if ((threadIdx.x & 63) == 0 && nleaf) atomicAdd(too_large + 1, nleaf);
    }
}

// end of a level: the next level's slots end where the counter stands
__global__ void k_bvh_mark(const uint32_t* __restrict__ counter, uint32_t* __restrict__ level_base, int level) {
    if (threadIdx.x == 0) level_base[level + 2] = *counter;
}

// 7. records in leaf order at each primitive's rank; leaf -> source index per type for the host
__global__ __launch_bounds__(256) void k_bvh_scatter(Prims P, uint32_t n, const uint32_t* __restrict__ vals,
                                                     const TypeCount* __restrict__ ranks, const float4* __restrict__ tri_nrm,
                                                     const float2* __restrict__ tri_uv, const float2* __restrict__ sph_map,
                                                     float4* __restrict__ geo_tris,
                                                     float4* __restrict__ geo_sph, float4* __restrict__ geo_quads,
                                                     float4* __restrict__ out_nrm, float2* __restrict__ out_uv,
                                                     float2* __restrict__ out_map,
                                                     uint32_t* __restrict__ src_tri, uint32_t* __restrict__ src_sph,
                                                     uint32_t* __restrict__ src_quad) {
    for (uint32_t s = blockIdx.x * 256u + threadIdx.x; s < n; s += gridDim.x * 256u) {
        const uint32_t p = vals[s];
        const uint32_t t = P.type_of(p);
        const uint32_t r = ranks[s].c[t];
        if (t == CRT_PRIM_TRI) {
            const uint32_t i = p - P.nq - P.ns;
            for (int k = 0; k < 3; ++k) geo_tris[3 * (size_t)r + k] = P.tris[3 * (size_t)i + k];
            if (out_nrm) for (int k = 0; k < 3; ++k) out_nrm[3 * (size_t)r + k] = tri_nrm[3 * (size_t)i + k];
            if (out_uv) for (int k = 0; k < 3; ++k) out_uv[3 * (size_t)r + k] = tri_uv[3 * (size_t)i + k];
            src_tri[r] = i;
        } else if (t == CRT_PRIM_SPHERE) {
            const uint32_t i = p - P.nq;
            geo_sph[2 * (size_t)r] = P.spheres[2 * (size_t)i];
            geo_sph[2 * (size_t)r + 1] = P.spheres[2 * (size_t)i + 1];
            for (int k = 0; k < 3; ++k) out_map[3 * (size_t)r + k] = sph_map[3 * (size_t)i + k];
            src_sph[r] = i;
        } else {
            for (int k = 0; k < 5; ++k) geo_quads[5 * (size_t)r + k] = P.quads[5 * (size_t)p + k];
            src_quad[r] = p;
        }
    }
}

__global__ void k_fill_u32_gb(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

uint32_t grid_for(uint64_t n, uint32_t cap = 4096) {
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(cap, (n + 255) / 256));
}

template <class T> T* dptr(crt::DevBuf& b) { return (T*)b.p; }

}  // namespace

namespace crt {

#define GB_LAUNCH(call, what)                                                        \
    do {                                                                             \
        call;                                                                        \
        hipError_t e_ = hipGetLastError();                                           \
        if (e_ != hipSuccess) return set_error(CRT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Builds a one-level scene's BVH and records on the device into `sc` (its geometry allocation,
// per-triangle / per-sphere side arrays, light table, host id maps and statistics). The caller
// holds the context's lock and checks the budget.
int gpu_build_scene(Context& C, const crt_scene_desc* d, uint32_t leaf, uint32_t width, DeviceScene& sc, std::string& err) {
    auto t0 = std::chrono::steady_clock::now();
    HostScene& H = sc.host;
    int rc = prepare_scene_tables(d, leaf, width, H, err);
    if (rc) return rc;
    SourceRecords src;
    unsigned hw = std::thread::hardware_concurrency();
    if ((rc = flatten_sources(d, H.shape_map_base, (int)std::min(16u, hw ? hw : 1u), src, err))) return rc;
    const uint32_t nq = (uint32_t)(src.quads.size() / 5), ns = (uint32_t)(src.spheres.size() / 2),
                   nt = (uint32_t)(src.tris.size() / 3);
    const uint64_t n64 = (uint64_t)nq + ns + nt;
    if (n64 < 2) { err = "the device build needs at least two primitives"; return CRT_ERR_INVALID_ARGUMENT; }
    if (n64 > (1ull << 29)) { err = "too many primitives for one scene"; return CRT_ERR_INVALID_SCENE; }
    const uint32_t n = (uint32_t)n64;
    hipStream_t s = C.stream;
    Context* ctx = &C;
    // source records on the device
    DevBuf squads, sqbox, ssph, stris, snrm, suv, smap;
    auto up = [&](DevBuf& b, const void* h, size_t bytes) -> int {
        if (!bytes) return CRT_OK;
        int r = b.alloc(bytes, ctx);
        if (r) return r;
        HIP_TRY(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, s));
        return CRT_OK;
    };
    if ((rc = up(squads, src.quads.data(), src.quads.size() * sizeof(F4)))) return rc;
    if ((rc = up(sqbox, src.quad_boxes.data(), src.quad_boxes.size() * sizeof(F4)))) return rc;
    if ((rc = up(ssph, src.spheres.data(), src.spheres.size() * sizeof(F4)))) return rc;
    if ((rc = up(smap, src.sphere_map.data(), src.sphere_map.size() * sizeof(F2)))) return rc;
    if ((rc = up(stris, src.tris.data(), src.tris.size() * sizeof(F4)))) return rc;
    if ((rc = up(snrm, src.tri_normals.data(), src.tri_normals.size() * sizeof(F4)))) return rc;
    if ((rc = up(suv, src.tri_uvs.data(), src.tri_uvs.size() * sizeof(F2)))) return rc;
    Prims P{dptr<float4>(squads), dptr<float4>(sqbox), dptr<float4>(ssph), dptr<float4>(stris), nq, ns, nt};

    // work buffers
    DevBuf plo, phi, keys[2], vals[2], cb, cl, cr, parent, lparent, nlo, nhi, first, count, tmask, visits, tflags, ranks,
        slot_b, level_base, counters, sah, nodes_tmp, tmp;
    const size_t ni = std::max<uint32_t>(n - 1, 1);
    if ((rc = plo.alloc((size_t)n * 16, ctx)) || (rc = phi.alloc((size_t)n * 16, ctx)) ||
        (rc = keys[0].alloc((size_t)n * 8, ctx)) || (rc = keys[1].alloc((size_t)n * 8, ctx)) ||
        (rc = vals[0].alloc((size_t)n * 4, ctx)) || (rc = vals[1].alloc((size_t)n * 4, ctx)) || (rc = cb.alloc(6 * 4, ctx)) ||
        (rc = cl.alloc(ni * 4, ctx)) || (rc = cr.alloc(ni * 4, ctx)) || (rc = parent.alloc(ni * 4, ctx)) ||
        (rc = lparent.alloc((size_t)n * 4, ctx)) || (rc = nlo.alloc(ni * 16, ctx)) || (rc = nhi.alloc(ni * 16, ctx)) ||
        (rc = first.alloc(ni * 4, ctx)) || (rc = count.alloc(ni * 4, ctx)) || (rc = tmask.alloc(ni * 4, ctx)) ||
        (rc = visits.alloc(ni * 4, ctx)) || (rc = tflags.alloc((size_t)n * sizeof(TypeCount), ctx)) ||
        (rc = ranks.alloc((size_t)n * sizeof(TypeCount), ctx)) || (rc = slot_b.alloc(ni * 4, ctx)) ||
        (rc = level_base.alloc((kMaxLevels + 2) * 4, ctx)) || (rc = counters.alloc(4 * 4, ctx)) || (rc = sah.alloc(8, ctx)))
        return rc;
    const uint32_t per = width == 4 ? 8u : 4u;   // float4 per node
    if ((rc = nodes_tmp.alloc(ni * per * 16, ctx))) return rc;
    // temporary storage of the sort and the scan
    size_t sort_bytes = 0, scan_bytes = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, dptr<uint64_t>(keys[0]), dptr<uint64_t>(keys[1]),
                                               dptr<uint32_t>(vals[0]), dptr<uint32_t>(vals[1]), (int)n, 0, 63, s));
    TypeCount zero{{0u, 0u, 0u}};
    HIP_TRY(hipcub::DeviceScan::ExclusiveScan(nullptr, scan_bytes, dptr<TypeCount>(tflags), dptr<TypeCount>(ranks),
                                              TypeCountAdd(), zero, (int)n, s));
    if ((rc = tmp.alloc(std::max<size_t>(std::max(sort_bytes, scan_bytes), 16), ctx))) return rc;

    // 1-2. boxes, centroid bounds, Morton codes, sort
    GB_LAUNCH(hipLaunchKernelGGL(k_fill_u32_gb, dim3(1), dim3(64), 0, s, dptr<uint32_t>(cb), 3u, 0xffffffffu), "fill");
    GB_LAUNCH(hipLaunchKernelGGL(k_fill_u32_gb, dim3(1), dim3(64), 0, s, dptr<uint32_t>(cb) + 3, 3u, 0u), "fill");
    GB_LAUNCH(hipLaunchKernelGGL(k_bvh_bounds, dim3(grid_for(n, 2048)), dim3(256), 0, s, P, n, dptr<float4>(plo),
                                 dptr<float4>(phi), dptr<uint32_t>(cb)), "bounds");
    GB_LAUNCH(hipLaunchKernelGGL(k_bvh_morton, dim3(grid_for(n)), dim3(256), 0, s, dptr<float4>(plo), dptr<float4>(phi),
                                 dptr<uint32_t>(cb), n, dptr<uint64_t>(keys[0]), dptr<uint32_t>(vals[0])), "morton");
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp.p, sort_bytes, dptr<uint64_t>(keys[0]), dptr<uint64_t>(keys[1]),
                                               dptr<uint32_t>(vals[0]), dptr<uint32_t>(vals[1]), (int)n, 0, 63, s));
    const uint64_t* skeys = dptr<uint64_t>(keys[1]);
    const uint32_t* svals = dptr<uint32_t>(vals[1]);
    // 3-5. tree, boxes, ranks
    GB_LAUNCH(hipLaunchKernelGGL(k_bvh_karras, dim3(grid_for(n)), dim3(256), 0, s, skeys, (int)n, dptr<int32_t>(cl),
                                 dptr<int32_t>(cr), dptr<int32_t>(parent), dptr<int32_t>(lparent), dptr<uint32_t>(first),
                                 dptr<uint32_t>(count)), "karras");
    GB_LAUNCH(hipLaunchKernelGGL(k_fill_u32_gb, dim3(grid_for(ni)), dim3(256), 0, s, dptr<uint32_t>(visits), (uint32_t)ni, 0u),
              "fill");
    GB_LAUNCH(hipLaunchKernelGGL(k_bvh_refit, dim3(grid_for(n)), dim3(256), 0, s, (int)n, svals, dptr<float4>(plo),
                                 dptr<float4>(phi), dptr<int32_t>(cl), dptr<int32_t>(cr), dptr<int32_t>(parent),
                                 dptr<int32_t>(lparent), dptr<float4>(nlo), dptr<float4>(nhi), dptr<uint32_t>(tmask),
                                 dptr<uint32_t>(visits)), "refit");
    GB_LAUNCH(hipLaunchKernelGGL(k_bvh_typeflags, dim3(grid_for(n)), dim3(256), 0, s, svals, P, n, dptr<TypeCount>(tflags)),
              "type flags");
    HIP_TRY(hipcub::DeviceScan::ExclusiveScan(tmp.p, scan_bytes, dptr<TypeCount>(tflags), dptr<TypeCount>(ranks),
                                              TypeCountAdd(), zero, (int)n, s));
    // root box (for the SAH statistic) and the collapse's start: slot 0 = binary node 0
    float4 root_lo, root_hi;
    HIP_TRY(hipMemcpyAsync(&root_lo, dptr<float4>(nlo), 16, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&root_hi, dptr<float4>(nhi), 16, hipMemcpyDeviceToHost, s));
    const uint32_t init_base[3] = {0u, 1u, 1u};
    HIP_TRY(hipMemcpyAsync(dptr<uint32_t>(level_base), init_base, sizeof(init_base), hipMemcpyHostToDevice, s));
    const uint32_t init_cnt[4] = {1u, 0u, 0u, 0u};   // node counter, too-large flag, leaves
    HIP_TRY(hipMemcpyAsync(dptr<uint32_t>(counters), init_cnt, sizeof(init_cnt), hipMemcpyHostToDevice, s));
    const int32_t root_b = 0;
    HIP_TRY(hipMemcpyAsync(dptr<int32_t>(slot_b), &root_b, 4, hipMemcpyHostToDevice, s));
    const double zero_d = 0.0;
    HIP_TRY(hipMemcpyAsync(sah.p, &zero_d, 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    const float dx = root_hi.x - root_lo.x, dy = root_hi.y - root_lo.y, dz = root_hi.z - root_lo.z;
    const float root_area = std::max(2.0f * (dx * dy + dy * dz + dz * dx), 1e-30f);
    Tree T{dptr<int32_t>(cl), dptr<int32_t>(cr), dptr<float4>(nlo), dptr<float4>(nhi), dptr<uint32_t>(first),
           dptr<uint32_t>(count), dptr<uint32_t>(tmask), svals, dptr<float4>(plo), dptr<float4>(phi), dptr<TypeCount>(ranks),
           P, std::max(1u, std::min(leaf, CRT_LEAF_MAX_COUNT))};
    // 6. collapse, level by level; the level bases are read back every few levels to stop
    std::vector<uint32_t> lb(kMaxLevels + 2, 0);
    int levels = 0;
    const uint32_t cgrid = grid_for(n / 2 + 1, 2048);
    for (int L = 0; L < kMaxLevels; ++L) {
        if (width == 4)
            GB_LAUNCH(hipLaunchKernelGGL(k_bvh_collapse<4>, dim3(cgrid), dim3(256), 0, s, T, dptr<uint32_t>(level_base), L,
                                         dptr<int32_t>(slot_b), dptr<uint32_t>(counters), dptr<float4>(nodes_tmp),
                                         dptr<double>(sah), 1.0f / root_area, dptr<uint32_t>(counters) + 1), "collapse");
        else
            GB_LAUNCH(hipLaunchKernelGGL(k_bvh_collapse<2>, dim3(cgrid), dim3(256), 0, s, T, dptr<uint32_t>(level_base), L,
                                         dptr<int32_t>(slot_b), dptr<uint32_t>(counters), dptr<float4>(nodes_tmp),
                                         dptr<double>(sah), 1.0f / root_area, dptr<uint32_t>(counters) + 1), "collapse");
        GB_LAUNCH(hipLaunchKernelGGL(k_bvh_mark, dim3(1), dim3(64), 0, s, dptr<uint32_t>(counters), dptr<uint32_t>(level_base), L),
                  "mark");
        if ((L & 7) == 7 || L == kMaxLevels - 1) {
            HIP_TRY(hipMemcpyAsync(lb.data(), level_base.p, lb.size() * 4, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            int done = -1;
            for (int k = 0; k <= L + 1; ++k) if (lb[k + 1] == lb[k]) { done = k; break; }
            if (done >= 0) { levels = done; break; }
        }
    }
    if (levels == 0) { err = "BVH deeper than the device build's level limit"; return CRT_ERR_INTERNAL; }
    uint32_t cnt_h[4];
    double sah_h = 0.0;
    HIP_TRY(hipMemcpyAsync(cnt_h, counters.p, sizeof(cnt_h), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&sah_h, sah.p, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (cnt_h[1]) { err = "more than 2^26 primitive records of one type"; return CRT_ERR_INVALID_SCENE; }
    const uint32_t nnodes = cnt_h[0];

    // 7. the scene's geometry allocation in the traversal's layout, records at their ranks
    const uint64_t node_bytes = (uint64_t)nnodes * per * 16;
    const uint64_t tri_off = HostScene::align256(node_bytes);
    const uint64_t inst_off = HostScene::align256(tri_off + (uint64_t)nt * 48);
    const uint64_t sph_off = HostScene::align256(inst_off);
    const uint64_t quad_off = HostScene::align256(sph_off + (uint64_t)ns * 32);
    const uint64_t geo_bytes = quad_off + (uint64_t)nq * 80;
    if (geo_bytes > kMaxGeometryBytes) {
        err = "BVH nodes and primitive records take " + std::to_string(geo_bytes) + " bytes; one scene holds at most 4 GiB of them";
        return CRT_ERR_INVALID_SCENE;
    }
    if ((rc = sc.geo.alloc(geo_bytes + 256, ctx))) return rc;
    char* geo = (char*)sc.geo.p;
    HIP_TRY(hipMemcpyAsync(geo, nodes_tmp.p, node_bytes, hipMemcpyDeviceToDevice, s));
    if (nt && !src.tri_normals.empty() && (rc = sc.tri_normals.alloc((size_t)nt * 48, ctx))) return rc;
    if (nt && !src.tri_uvs.empty() && (rc = sc.tri_uvs.alloc((size_t)nt * 24, ctx))) return rc;
    if (ns && (rc = sc.sphere_map.alloc((size_t)ns * 24, ctx))) return rc;
    DevBuf srcmap;
    if ((rc = srcmap.alloc((size_t)(nt + ns + nq) * 4 + 16, ctx))) return rc;
    uint32_t* src_tri = dptr<uint32_t>(srcmap);
    uint32_t* src_sph = src_tri + nt;
    uint32_t* src_quad = src_sph + ns;
    GB_LAUNCH(hipLaunchKernelGGL(k_bvh_scatter, dim3(grid_for(n)), dim3(256), 0, s, P, n, svals, dptr<TypeCount>(ranks),
                                 dptr<float4>(snrm), dptr<float2>(suv), dptr<float2>(smap),
                                 (float4*)(geo + tri_off), (float4*)(geo + sph_off), (float4*)(geo + quad_off),
                                 dptr<float4>(sc.tri_normals), dptr<float2>(sc.tri_uvs), dptr<float2>(sc.sphere_map),
                                 src_tri, src_sph, src_quad), "scatter");
    std::vector<uint32_t> srch((size_t)nt + ns + nq);
    if (!srch.empty()) HIP_TRY(hipMemcpyAsync(srch.data(), srcmap.p, srch.size() * 4, hipMemcpyDeviceToHost, s));
    if ((rc = (sc.mats.alloc(H.mats.size() * 16, ctx)))) return rc;
    if (!H.mats.empty()) HIP_TRY(hipMemcpyAsync(sc.mats.p, H.mats.data(), H.mats.size() * 16, hipMemcpyHostToDevice, s));
    if (!H.texs.empty()) {
        if ((rc = sc.texs.alloc(H.texs.size() * 16, ctx))) return rc;
        HIP_TRY(hipMemcpyAsync(sc.texs.p, H.texs.data(), H.texs.size() * 16, hipMemcpyHostToDevice, s));
    }
    if (!H.texels.empty()) {
        if ((rc = sc.texels.alloc(H.texels.size(), ctx))) return rc;
        HIP_TRY(hipMemcpyAsync(sc.texels.p, H.texels.data(), H.texels.size(), hipMemcpyHostToDevice, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    // host maps: leaf order -> global id
    sc.tri_gid.resize(nt);
    for (uint32_t r = 0; r < nt; ++r) sc.tri_gid[r] = src.tri_gid0 + srch[r];
    H.sphere_aux.resize(ns);
    for (uint32_t r = 0; r < ns; ++r) H.sphere_aux[r] = src.sphere_aux[srch[(size_t)nt + r]];
    sc.quad_gid.resize(nq);
    for (uint32_t r = 0; r < nq; ++r) sc.quad_gid[r] = src.quad_gid[srch[(size_t)nt + ns + r]];
    sc.inst_gid.clear();
    // light table: each light's source record -> its leaf-order record
    if (!src.light_src.empty()) {
        std::vector<uint32_t> sph_slot(ns), quad_slot(nq), table(src.light_src.size());
        for (uint32_t r = 0; r < ns; ++r) sph_slot[srch[(size_t)nt + r]] = r;
        for (uint32_t r = 0; r < nq; ++r) quad_slot[srch[(size_t)nt + ns + r]] = r;
        for (size_t i = 0; i < table.size(); ++i) {
            const uint32_t s0 = src.light_src[i];
            table[i] = (s0 & CRT_LIGHT_SPHERE) ? (CRT_LIGHT_SPHERE | sph_slot[s0 & ~CRT_LIGHT_SPHERE]) : quad_slot[s0];
        }
        if ((rc = sc.lights.alloc(table.size() * 4, ctx))) return rc;
        HIP_TRY(hipMemcpyAsync(sc.lights.p, table.data(), table.size() * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    // device scene record
    crt_dev_scene& ds = sc.dev;
    std::memset(&ds, 0, sizeof(ds));
    ds.nodes = geo; ds.tris = geo + tri_off; ds.insts = geo + inst_off; ds.spheres = geo + sph_off; ds.quads = geo + quad_off;
    ds.tri_normals = sc.tri_normals.p; ds.tri_uvs = sc.tri_uvs.p; ds.sphere_map = sc.sphere_map.p;
    ds.mats = sc.mats.p; ds.texs = sc.texs.p; ds.texels = (const uint8_t*)sc.texels.p;
    ds.lights = (const uint32_t*)sc.lights.p;
    ds.num_lights = src.num_lights; ds.root = 0; ds.num_quads = d->num_quads; ds.bvh_width = width;
    for (int k = 0; k < 3; ++k) ds.background[k] = H.background[k];
    H.num_lights = src.num_lights; H.num_spheres = ns; H.num_quads = d->num_quads; H.num_instances = 0;
    H.num_shapes = d->num_shapes;
    H.num_quad_records = nq;
    H.num_world_tris = nt;
    H.root = 0;
    H.stats.nodes = nnodes;
    H.num_node_records = nnodes;
    H.num_tri_records = nt;
    H.stats.leaves = cnt_h[2];
    H.stats.prims = n;
    H.stats.depth = (uint32_t)levels;
    H.stats.sah_cost = sah_h;
    H.device_built = true;
    H.stats.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    sc.bytes = sc.geo.bytes + sc.tri_normals.bytes + sc.tri_uvs.bytes + sc.sphere_map.bytes + sc.mats.bytes +
               sc.texs.bytes + sc.texels.bytes + sc.lights.bytes;
    // the host copies of the tables are not needed after the upload
    H.texels.clear(); H.texels.shrink_to_fit();
    return CRT_OK;
}

}  // namespace crt
