// kernels.hip - gfx950 wavefront path-tracing kernels.
//
// One pass of the hot path (core/src/gpu.cu:761-802 per pixel, core/src/cameras/camera.cuh:149-264
// per bounce) is split into stages that run over every live path of every request in the pass:
//   k_regen     camera rays into free path slots   (render kernel sample loop + get_ray)
//   k_extend    BVH traversal, closest hit          (hittable_list/bvh_node/aabb/primitive ::hit)
//   k_shade     emission, scatter, light mixture    (ray_color body, materials, pdfs, textures)
//               + compaction of the survivors (wave ballot, block prefix scan, one atomic per
//               block) into the other state buffer; finished paths write their radiance by path id
// The stages repeat over a fixed pool of path slots until every path of the pass has finished:
// slots freed by one iteration are refilled with new camera paths by the next, so each launch
// works on a full queue instead of the thinning tail of one generation of paths.
//   k_resolve   per-pixel sums in sample order      (pixel_color += ray_color)
//   k_tonemap   NaN->0, /spp, sqrt, clamp, *255.99  (prepare_pixel_color + quantize)
// MFMA is not used: nothing here is a dense contraction. The traversal is bound by node/primitive
// fetches (HBM / L2 / MALL) and VALU; layouts are SoA float4 so every fetch is a dwordx4.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cortex_rt.h"
#include "crt_layout.h"
#include "crt_kernels.h"

namespace {

// ------------------------------------------------------------------ small vector helpers
struct V3 { float x, y, z; };
__device__ __forceinline__ V3 v3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ V3 v3(float4 a) { return {a.x, a.y, a.z}; }
__device__ __forceinline__ V3 v3p(const float* p) { return {p[0], p[1], p[2]}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ float len2(V3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
__device__ __forceinline__ float vlen(V3 v) { return sqrtf(len2(v)); }
__device__ __forceinline__ V3 unit(V3 v) { float l = vlen(v); return {v.x / l, v.y / l, v.z / l}; }
__device__ __forceinline__ float comp(V3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float bitsf(uint32_t u) { return __uint_as_float(u); }

template <class T> __device__ __forceinline__ T ldg(const T* p) { return *p; }

// ------------------------------------------------------------------ random streams
// Thrust minstd_rand step with Carta's reduction for m = 2^31 - 1 (exact).
__device__ __forceinline__ uint32_t minstd_next(uint32_t& x) {
    uint64_t p = (uint64_t)x * 48271u;
    uint32_t r = (uint32_t)(p & 0x7fffffffu) + (uint32_t)(p >> 31);
    if (r >= 0x7fffffffu) r -= 0x7fffffffu;
    x = r;
    return r;
}
// uniform_real_distribution<float>(0,1): float(x-1) / 2^31 (the float denominator rounds to 2^31)
__device__ __forceinline__ float minstd_uniform(uint32_t& x) {
    float r = (float)(minstd_next(x) - 1u);
    return r / 2147483648.0f;
}
struct Pcg { uint64_t state, inc; };
__device__ __forceinline__ uint32_t pcg_next(Pcg& g) {
    uint64_t old = g.state;
    g.state = old * 6364136223846793005ULL + g.inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ Pcg pcg_keyed(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce) {
    uint64_t key = ((uint64_t)pixel << 32) | (uint64_t)sample;
    uint64_t initstate = splitmix64(seed ^ splitmix64(key));
    Pcg g;
    g.state = 0u;
    g.inc = ((uint64_t)bounce << 1u) | 1u;
    pcg_next(g);
    g.state += initstate;
    pcg_next(g);
    return g;
}

// One draw source. MODE is CRT_RNG_PCG or CRT_RNG_MINSTD (a template constant: each kernel
// instance is branch-free) or CRT_RNG_DYNAMIC (the mode is read per path from its request, for
// launches that mix requests of both kinds).
template <int MODE> struct Rng {
    uint32_t ms;
    Pcg pcg;
    uint32_t dyn;   // CRT_RNG_DYNAMIC: this path's mode
    __device__ __forceinline__ bool minstd() const {
        return MODE == CRT_RNG_MINSTD || (MODE == CRT_RNG_DYNAMIC && dyn == CRT_RNG_MINSTD);
    }
    __device__ __forceinline__ float uniform() {
        if (minstd()) return minstd_uniform(ms);
        return (float)(pcg_next(pcg) >> 8) * (1.0f / 16777216.0f);
    }
    __device__ __forceinline__ void skip() {
        if (minstd()) minstd_next(ms); else pcg_next(pcg);
    }
};

// ------------------------------------------------------------------ global loads
// Scene arrays are reached through pointers read from the scene table, so the compiler cannot
// tell they are global memory and emits flat loads (which also wait on the LDS counter the
// traversal stack uses). These loads go through address-space-1 vector pointers: global_load.
typedef float crt_v4f __attribute__((ext_vector_type(4)));
typedef int crt_v4i __attribute__((ext_vector_type(4)));
typedef unsigned crt_v2u __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float4 ldg(const float4* p) {
    crt_v4f v = *(__attribute__((address_space(1))) const crt_v4f*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
// Load at base + a 32-bit byte offset: with a wave-uniform base the address is one scalar pair
// plus one VGPR per load (the saddr form), not a 64-bit VGPR address per load.
__device__ __forceinline__ float4 ldg_at(const float4* base, uint32_t byte_off) {
    typedef __attribute__((address_space(1))) const char gchar;
    crt_v4f v = *(__attribute__((address_space(1))) const crt_v4f*)((gchar*)base + byte_off);
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ int4 ldg_i4(const float4* p) {
    crt_v4i v = *(__attribute__((address_space(1))) const crt_v4i*)p;
    return make_int4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint2 ldg(const uint2* p) {
    crt_v2u v = *(__attribute__((address_space(1))) const crt_v2u*)p;
    return make_uint2(v.x, v.y);
}
typedef unsigned crt_v4u __attribute__((ext_vector_type(4)));
typedef float crt_v2f __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint4 ldg(const uint4* p) {
    crt_v4u v = *(__attribute__((address_space(1))) const crt_v4u*)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float2 ldg(const float2* p) {
    crt_v2f v = *(__attribute__((address_space(1))) const crt_v2f*)p;
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ uint32_t ldg(const uint32_t* p) { return *(__attribute__((address_space(1))) const uint32_t*)p; }
__device__ __forceinline__ uint8_t ldg(const uint8_t* p) { return *(__attribute__((address_space(1))) const uint8_t*)p; }
// stores through pointers read from request records (same reason: global, not flat)
__device__ __forceinline__ void stg(uint32_t* p, uint32_t v) { *(__attribute__((address_space(1))) uint32_t*)p = v; }
__device__ __forceinline__ void stg(float4* p, float4 v) {
    crt_v4f w = {v.x, v.y, v.z, v.w};
    *(__attribute__((address_space(1))) crt_v4f*)p = w;
}

// ------------------------------------------------------------------ traversal stack
// Short per-lane stack: RS entries in VGPRs (static shifts, no dynamic register indexing; one entry
// is used: the top), spilling to an LS-entry ring per lane in LDS, which in turn spills its oldest
// entries to a per-thread overflow area in global memory. A pop returns the register top at once
// and reads the entry below it from LDS while the popped node is fetched.
// Overflow area of one thread: the launch-wide base (kernel argument, scalar) and the thread's
// offset in entries (one VGPR instead of a 64-bit per-lane pointer).
struct GSlot { uint32_t* base; uint32_t off; };

// The LDS ring and the overflow area are held as address-space-qualified pointers, so the two
// refill sources stay a ds_read and a global_load: merged through one generic pointer they become
// a flat load that waits on both memory counters at every pop.
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(1))) uint32_t glb_u32;

template <int BLOCK, int RS, int LS, int GS> struct Stack {
    uint32_t r[RS];
    int rc, lbase, lcnt, gsp;
    lds_u32* lds;      // column of this lane: lds[e * BLOCK]
    glb_u32* gbase;    // overflow area of the launch (scalar); this thread's entries start at goff
    uint32_t goff;
    bool overflow;
    __device__ __forceinline__ void init(uint32_t* lds_col, GSlot g) {
        rc = 0; lbase = 0; lcnt = 0; gsp = 0; overflow = false;
        lds = (lds_u32*)lds_col;
        gbase = (glb_u32*)g.base;
        goff = g.off;
    }
    __device__ __forceinline__ int depth() const { return rc + lcnt + gsp; }
    __device__ __forceinline__ bool empty() const { return rc == 0; }
    __device__ __forceinline__ void spill(uint32_t v) {
        if (lcnt == LS) {
            if (gsp < GS) gbase[goff + gsp++] = lds[lbase * BLOCK]; else overflow = true;
            lbase = (lbase + 1) & (LS - 1);
            --lcnt;
        }
        lds[((lbase + lcnt) & (LS - 1)) * BLOCK] = v;
        ++lcnt;
    }
    __device__ __forceinline__ uint32_t refill() {
        if (lcnt > 0) { --lcnt; return lds[((lbase + lcnt) & (LS - 1)) * BLOCK]; }
        return gbase[goff + --gsp];
    }
    __device__ __forceinline__ void push(uint32_t v) {
        if (rc == RS) spill(r[RS - 1]); else ++rc;
#pragma unroll
        for (int k = RS - 1; k > 0; --k) r[k] = r[k - 1];
        r[0] = v;
    }
    __device__ __forceinline__ uint32_t pop() {
        uint32_t v = r[0];
#pragma unroll
        for (int k = 0; k < RS - 1; ++k) r[k] = r[k + 1];
        if (lcnt > 0 || gsp > 0) r[RS - 1] = refill(); else --rc;
        return v;
    }
};

// ------------------------------------------------------------------ primitive tests
struct TriPre { uint32_t k; float Sx, Sy, Sz; };   // k = kx | ky << 2 | kz << 4
__device__ __forceinline__ TriPre tri_pre(V3 d) {
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    float dz = comp(d, kz);
    if (dz < 0.0f) { int t = kx; kx = ky; ky = t; }
    TriPre p;
    p.k = (uint32_t)kx | ((uint32_t)ky << 2) | ((uint32_t)kz << 4);
    p.Sx = comp(d, kx) / dz; p.Sy = comp(d, ky) / dz; p.Sz = 1.0f / dz;
    return p;
}

// The watertight test evaluated in double from the same float inputs (ray direction d, origin o,
// vertices): used when a float edge function is too close to zero for its sign to be certain, so a
// ray near an edge is put on the same side of it as an FP64 evaluation puts it.
__device__ __forceinline__ bool tri_hit_f64(V3 o, V3 d, float4 a, float4 b, float4 c, float tmin, float tmax, float& t, float& bu,
                                         float& bv) {
#pragma clang fp contract(off)
    const double dx = d.x, dy = d.y, dz = d.z;
    const double ax = fabs(dx), ay = fabs(dy), az = fabs(dz);
    int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    auto cd = [](double x, double y, double z, int k) { return k == 0 ? x : (k == 1 ? y : z); };
    const double dkz = cd(dx, dy, dz, kz);
    if (dkz < 0.0) { int tk = kx; kx = ky; ky = tk; }
    const double Sx = cd(dx, dy, dz, kx) / dkz, Sy = cd(dx, dy, dz, ky) / dkz, Sz = 1.0 / dkz;
    const double Aa[3] = {(double)a.x - (double)o.x, (double)a.y - (double)o.y, (double)a.z - (double)o.z};
    const double Ba[3] = {(double)b.x - (double)o.x, (double)b.y - (double)o.y, (double)b.z - (double)o.z};
    const double Ca[3] = {(double)c.x - (double)o.x, (double)c.y - (double)o.y, (double)c.z - (double)o.z};
    const double Ax = Aa[kx] - Sx * Aa[kz], Ay = Aa[ky] - Sy * Aa[kz];
    const double Bx = Ba[kx] - Sx * Ba[kz], By = Ba[ky] - Sy * Ba[kz];
    const double Cx = Ca[kx] - Sx * Ca[kz], Cy = Ca[ky] - Sy * Ca[kz];
    const double U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if ((U < 0.0 || V < 0.0 || W < 0.0) && (U > 0.0 || V > 0.0 || W > 0.0)) return false;
    const double det = U + V + W;
    if (det == 0.0) return false;
    const double T = U * (Sz * Aa[kz]) + V * (Sz * Ba[kz]) + W * (Sz * Ca[kz]);
    const double inv = 1.0 / det;
    const double tt = T * inv;
    if (!(tt >= (double)tmin && tt <= (double)tmax)) return false;
    t = (float)tt; bu = (float)(V * inv); bv = (float)(W * inv);
    return true;
}

// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2(1) 2013), two-sided. The edge
// functions are computed without FMA contraction so a shared edge gives exactly negated values
// in both triangles.
//
// The edge functions carry a common bound E on their FP32 rounding error. With u = 2^-24, a sheared
// coordinate such as Cx = C[kx] - Sx * C[kz] is within 4u mCx of its exact value, mCx = |C[kx]| +
// |Sx * C[kz]| (C = vertex - origin and Sx rounded once each, the product and the difference once
// more); cancellation makes that large against Cx itself when the vertex is far along the ray. A
// product Cx * By is then within 4u (|Cx| mBy + mCx |By|) + 16u^2 mCx mBy of the exact one, and
// rounding the products and their difference adds 2u (|Cx By| + |Cy Bx|), so U = Cx By - Cy Bx is
// within 12u (aX mY + mX aY) + 32u^2 mX mY, where aX, aY are the largest |x|, |y| of the three
// sheared vertices and mX, mY the largest magnitudes m; the same holds for V and W. E takes 16u and
// 64u^2. An edge function larger than E has the sign of the exact value, and of a double
// evaluation of the same expression (some 2^29 times more accurate). The FP32 answer is therefore
// the double answer whenever every sign is certain, or when one is certainly negative and another
// certainly positive (a miss either way). Otherwise the ray passes within rounding distance of an
// edge or a vertex: EXACT traversals then redo the test in double (tri_hit_f64); the others report
// the test undecided (`amb`), which ends the query, and the lane re-traces it with an EXACT
// traversal outside its traversal loop.
template <bool EXACT>
__device__ __forceinline__ bool tri_hit(V3 o, V3 d, const TriPre& p, float4 a, float4 b, float4 c, float tmin, float tmax,
                                        float& t, float& bu, float& bv, bool& amb) {
#pragma clang fp contract(off)
    const int kx = (int)(p.k & 3u), ky = (int)((p.k >> 2) & 3u), kz = (int)(p.k >> 4);
    V3 A = v3(a) - o, B = v3(b) - o, C = v3(c) - o;
    float Akz = comp(A, kz), Bkz = comp(B, kz), Ckz = comp(C, kz);
    float Akx = comp(A, kx), Aky = comp(A, ky), Bkx = comp(B, kx), Bky = comp(B, ky), Ckx = comp(C, kx), Cky = comp(C, ky);
    float pAx = p.Sx * Akz, pAy = p.Sy * Akz, pBx = p.Sx * Bkz, pBy = p.Sy * Bkz, pCx = p.Sx * Ckz, pCy = p.Sy * Ckz;
    float Ax = Akx - pAx, Ay = Aky - pAy;
    float Bx = Bkx - pBx, By = Bky - pBy;
    float Cx = Ckx - pCx, Cy = Cky - pCy;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    const float mX = fmaxf(fabsf(Akx) + fabsf(pAx), fmaxf(fabsf(Bkx) + fabsf(pBx), fabsf(Ckx) + fabsf(pCx)));
    const float mY = fmaxf(fabsf(Aky) + fabsf(pAy), fmaxf(fabsf(Bky) + fabsf(pBy), fabsf(Cky) + fabsf(pCy)));
    const float aX = fmaxf(fabsf(Ax), fmaxf(fabsf(Bx), fabsf(Cx)));
    const float aY = fmaxf(fabsf(Ay), fmaxf(fabsf(By), fabsf(Cy)));
    const float E = (aX * mY + mX * aY) * 0x1p-20f + (mX * mY) * 0x1p-42f;
    // smallest and largest edge function: a certain miss when one is below -E and another above E;
    // every sign certain when all three are beyond E on one side
    const float lo = fminf(U, fminf(V, W)), hi = fmaxf(U, fmaxf(V, W));
    if (lo < -E && hi > E) return false;
    if (!(lo > E || hi < -E)) {
        if (EXACT) return tri_hit_f64(o, d, a, b, c, tmin, tmax, t, bu, bv);
        amb = true;
        return false;
    }
    float det = U + V + W;
    float T = U * (p.Sz * Akz) + V * (p.Sz * Bkz) + W * (p.Sz * Ckz);
    float inv = 1.0f / det;
    float tt = T * inv;
    if (!(tt >= tmin && tt <= tmax)) return false;
    t = tt; bu = V * inv; bv = W * inv;
    return true;
}

// Sphere: strict interval like sphere::hit (core/src/primitives/sphere.cuh:111-132), computed with
// the perpendicular-distance discriminant and the stable root pair.
__device__ __forceinline__ bool sphere_hit(V3 o, V3 d, float4 s, float tmin, float tmax, float& t) {
    V3 oc = o - v3(s);
    float a = len2(d);
    float hb = dot(oc, d);
    float cc = len2(oc) - s.w * s.w;
    V3 l = oc - (hb / a) * d;
    float disc = a * (s.w * s.w - len2(l));
    if (disc < 0.0f) return false;
    float sq = sqrtf(disc);
    float q = hb >= 0.0f ? -(hb + sq) : -(hb - sq);
    float r0, r1;
    if (q == 0.0f) { r0 = r1 = -hb / a; }
    else { r0 = cc / q; r1 = q / a; }
    if (r0 > r1) { float x = r0; r0 = r1; r1 = x; }
    if (tmin < r0 && r0 < tmax) { t = r0; return true; }
    if (tmin < r1 && r1 < tmax) { t = r1; return true; }
    return false;
}

// Quad / area light: quad::hit (core/src/primitives/quad.cuh:100-133) and directional_light::hit
// (core/src/lights/directional_light.cuh:76-118), inclusive interval. `key` = the quad's global id.
// Analytic shapes, whose records share the quad array (flags CRT_QF_DISK / CYLINDER / CONE). The ray
// enters the shape's frame as under the reference's scale / rotate / translate wrappers
// (local = A x + b, composed on the host; untransformed, local origin = origin - centre exactly).
__device__ __forceinline__ void shape_local_ray(const float4& q0, const float4& q2, const float4& q3, const float4& q4,
                                                V3 o, V3 d, V3& lo, V3& ld) {
    lo = v3(q0.x * o.x + q0.y * o.y + q0.z * o.z + q0.w, q2.x * o.x + q2.y * o.y + q2.z * o.z + q2.w,
            q3.x * o.x + q3.y * o.y + q3.z * o.z + q4.x);
    ld = v3(q0.x * d.x + q0.y * d.y + q0.z * d.z, q2.x * d.x + q2.y * d.y + q2.z * d.z, q3.x * d.x + q3.y * d.y + q3.z * d.z);
}

// disk::hit (core/src/primitives/disk.cuh:53-82): the plane y = 0 of the frame, radius r;
// cylinder::hit (cylinder.cuh:63-103): the open tube x^2 + z^2 = r^2, 0 <= y <= h, the near root
// taken when it lies in the interval, else the far one, and then rejected when outside the height
// (the far root is not tried again); cone::hit (cone.cuh:58-128): x^2 + z^2 = (r/h)^2 (h - y)^2,
// 0 <= y <= h, rejected when |discriminant| < 0.001. The interval is closed. Comparisons are
// written so that a NaN or infinite distance (a ray parallel to the disk or the axis) is no hit,
// where the reference's would accept some of them.
__device__ __forceinline__ bool shape_hit(V3 o, V3 d, const float4* q, uint32_t flags, float tmin, float tmax, float& t) {
    // the frame's rows are read again here, one at a time (they are in the cache), so the shape path
    // holds no more registers than the quad path it shares the traversal loop with
    V3 lo, ld;
    float4 row = ldg(q);
    lo.x = row.x * o.x + row.y * o.y + row.z * o.z + row.w;
    ld.x = row.x * d.x + row.y * d.y + row.z * d.z;
    row = ldg(q + 2);
    lo.y = row.x * o.x + row.y * o.y + row.z * o.z + row.w;
    ld.y = row.x * d.x + row.y * d.y + row.z * d.z;
    row = ldg(q + 3);
    const float bz = ldg(q + 4).x;
    lo.z = row.x * o.x + row.y * o.y + row.z * o.z + bz;
    ld.z = row.x * d.x + row.y * d.y + row.z * d.z;
    const float4 q1 = ldg(q + 1);
    const float r = q1.x, h = q1.y;
    if (flags & CRT_QF_DISK) {
        const float tt = (-lo.y) / ld.y;
        if (!(tmin <= tt && tt <= tmax)) return false;
        const float hx = lo.x + tt * ld.x, hz = lo.z + tt * ld.z;
        if (!(hx * hx + hz * hz <= r * r)) return false;
        t = tt;
        return true;
    }
    // one quadric for both: a = dx^2 + dz^2 - k dy^2, b = 2 x dx + 2 z dz + 2 k D dy,
    // c = x^2 + z^2 - (k D^2 + m r^2) with D = h - y, k = 0 and m = 1 for the cylinder, k = (r/h)^2
    // and m = 0 for the cone: each term is the reference's own expression (scaling by 2 and 4 and
    // adding zero are exact), the cone's threshold is 0 for the cylinder
    const bool cone = (flags & CRT_QF_CONE) != 0;
    const float D = h - lo.y;
    const float k = cone ? (r / h) * (r / h) : 0.0f;
    const float m = cone ? 0.0f : r * r;
    const float a = ld.x * ld.x + ld.z * ld.z - k * (ld.y * ld.y);
    const float b = 2.0f * lo.x * ld.x + 2.0f * lo.z * ld.z + 2.0f * k * D * ld.y;
    const float c = lo.x * lo.x + lo.z * lo.z - (k * (D * D) + m);
    const float disc = b * b - 4.0f * (a * c);
    if (fabsf(disc) < (cone ? 0.001f : 0.0f) || !(disc >= 0.0f)) return false;
    const float sq = sqrtf(disc);
    float root = (-b - sq) / (2.0f * a);
    if (!(tmin <= root && root <= tmax)) {
        root = (-b + sq) / (2.0f * a);
        if (!(tmin <= root && root <= tmax)) return false;
    }
    const float y = lo.y + root * ld.y;
    if (!(y >= 0.0f && y <= h)) return false;
    t = root;
    return true;
}

// SHAPES = false for records known to be quads or area lights (the light table).
template <bool SHAPES = true>
__device__ __forceinline__ bool quad_hit(V3 o, V3 d, const float4* q, float tmin, float tmax, bool camera,
                                         float& t, float& qa, float& qb, uint32_t& key) {
    float4 q0 = ldg(q), q1 = ldg(q + 1), q2 = ldg(q + 2), q3 = ldg(q + 3), q4 = ldg(q + 4);
    key = fbits(q3.w);
    if (SHAPES && (fbits(q1.w) & CRT_QF_SHAPE)) {
        qa = 0.0f; qb = 0.0f;
        return shape_hit(o, d, q, fbits(q1.w), tmin, tmax, t);
    }
    V3 n = v3(q3);
    float denom = dot(n, d);
    if (fabsf(denom) < 1e-8f) return false;
    float tt = (q0.w - dot(n, o)) / denom;
    if (!(tmin <= tt && tt <= tmax)) return false;
    V3 p = o + tt * d;
    V3 hp = p - v3(q0);
    V3 w = v3(q4);
    float alpha = dot(w, cross(hp, v3(q2)));
    float beta = dot(w, cross(v3(q1), hp));
    uint32_t flags = fbits(q1.w);
    if (flags & CRT_QF_LIGHT) {
        if (alpha < -0.5f || 0.5f < alpha || beta < -0.5f || 0.5f < beta) return false;
        if ((flags & CRT_QF_INVISIBLE) && camera) return false;
        qa = alpha + 0.5f; qb = beta + 0.5f;
    } else {
        if (alpha < 0.0f || 1.0f < alpha || beta < 0.0f || 1.0f < beta) return false;
        qa = alpha; qb = beta;
    }
    t = tt;
    return true;
}

// Closest hit with a deterministic tie rule: hits whose distances lie within a relative 2^-20 of each
// other tie, and the lower global primitive id (`key`) wins, whatever order the traversal finds them
// in (two triangles sharing an edge both report a ray through that edge, at distances a few ulps
// apart). A candidate is tested up to best.t * kTieHi (boxes too, so a tying primitive in a later
// box is still reached) and taken when it is nearer than best.t * kTieLo or ties with a lower key.
// With no hit yet the key is 0, which nothing undercuts: a hit must then be nearer than the ray's
// end by the relative margin. Both factors are exact in FP32.
#define CRT_TIE_HI 1.00000095367431640625f   // 1 + 2^-20
#define CRT_TIE_LO 0.99999904632568359375f   // 1 - 2^-20
struct HitRec { float t, u, v; uint32_t prim, inst, key; };
__device__ __forceinline__ bool hit_wins(float t, uint32_t key, const HitRec& b) {
    return t < b.t * CRT_TIE_LO || key < b.key;
}

// Traversal view of a scene: BVH nodes, triangle, instance, sphere and quad records live in one
// allocation (in that order, below 4 GiB), addressed by 32-bit byte offsets from the node array.
struct SceneView {
    const float4* nodes;
    uint32_t root, tri_off, inst_off, sph_off, quad_off;
    __device__ __forceinline__ const float4* quad(uint32_t i) const {
        return (const float4*)((const char*)nodes + quad_off + i * 80u);
    }
    __device__ __forceinline__ const float4* sphere(uint32_t i) const {
        return (const float4*)((const char*)nodes + sph_off + i * 32u);
    }
};

// Whether the closest hit h of ray (o, d) is an emitter seen from behind that the estimator passes
// through (core/src/cameras/camera.cuh:179-183: emitted alpha 0 starts a re-trace from t + 0.001):
// a one-sided area light's back face (directional_light.cuh:96-104) or a spot light's back face
// (diffuse_spot_light.cuh:41-44, which also adds white when the light is visible). Returns 0,
// CRT_PASSTHROUGH, or CRT_PASSTHROUGH | CRT_PASSTHROUGH_WHITE. The front-face test repeats the
// shade stage's arithmetic, so both stages see the same face.
__device__ __forceinline__ uint32_t passes_through(const SceneView& S, const HitRec& h, V3 o, V3 d) {
    if (h.prim == CRT_MISS) return 0u;
    const uint32_t type = CRT_HIT_TYPE(h.prim);
    if (type == CRT_PRIM_QUAD) {
        const float4* q = S.quad(CRT_HIT_INDEX(h.prim));
        const uint32_t fl = fbits(ldg(q + 1).w);
        return ((fl & CRT_QF_ONE_SIDED) && !(dot(d, v3(ldg(q + 3))) < 0.0f)) ? CRT_PASSTHROUGH : 0u;
    }
    if (type == CRT_PRIM_SPHERE) {
        const float4* sp = S.sphere(CRT_HIT_INDEX(h.prim));
        const uint32_t fl = fbits(ldg(sp + 1).z);
        if (!(fl & CRT_SF_SPOT)) return 0u;
        const float4 s = ldg(sp);
        const V3 p = o + h.t * d;
        const V3 outward = (p - v3(s)) / s.w;
        if (dot(d, outward) < 0.0f) return 0u;
        return CRT_PASSTHROUGH | ((fl & CRT_SF_INVISIBLE) ? 0u : CRT_PASSTHROUGH_WHITE);
    }
    return 0u;
}
__device__ __forceinline__ SceneView scene_view(const crt_dev_scene* sc) {
    SceneView v;
    v.nodes = (const float4*)sc->nodes;
    v.root = sc->root;
    const char* b = (const char*)sc->nodes;
    v.tri_off = (uint32_t)((const char*)sc->tris - b);
    v.inst_off = (uint32_t)((const char*)sc->insts - b);
    v.sph_off = (uint32_t)((const char*)sc->spheres - b);
    v.quad_off = (uint32_t)((const char*)sc->quads - b);
    return v;
}

struct TravCount { uint32_t nodes, leaves, tris, spheres, quads, insts; };

// Slab test of one child box on (lo - o) * inv_d, Ize's robust traversal: each slab distance is
// within a relative 4u of the exact one (u = 2^-24: the subtraction, the hardware reciprocal of at
// most 1 ulp, the product), so near <= far * (1 + 8u) plus the rounding of that product holds for
// every box the ray meets; CRT_BOX_FAR = 1 + 5 * 2^-23 covers it. A box entered at +inf is a miss:
// an empty child slot (+inf planes) gives near = +inf and, for a ray with nothing hit yet (tmax =
// FLT_MAX), far * CRT_BOX_FAR overflows to +inf as well, so the comparison alone would accept it
// and the traversal would follow the empty reference. (The four-wide test below needs no such
// term: its callers turn a hit at +inf into a miss.)
#define CRT_BOX_FAR 1.0000006f
__device__ __forceinline__ float box_near(float lx, float hx, float ly, float hy, float lz, float hz, V3 o, V3 id,
                                          float tmin, float tmax, bool& hit) {
    float ax = (lx - o.x) * id.x, bx = (hx - o.x) * id.x;
    float ay = (ly - o.y) * id.y, by = (hy - o.y) * id.y;
    float az = (lz - o.z) * id.z, bz = (hz - o.z) * id.z;
    float nr = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), tmin));
    float fr = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), tmax));
    hit = nr <= fr * CRT_BOX_FAR && nr < __builtin_inff();
    return nr;
}

// The four-wide slab test, one fused multiply-add per plane: t = p * inv_d - o * inv_d. The
// product o * inv_d is rounded once, which moves every plane distance by up to 2^-24 |o * inv_d| on
// top of the relative error the far factor covers; near planes subtract on = o * inv_d + e and far
// planes of = o * inv_d - e (e = 2^-22 |o * inv_d| per axis), so the test stays conservative. The
// near and far plane of each axis are chosen by the sign of the inverse direction (per ray): the
// fused form is monotone in the plane coordinate, so near <= far holds after rounding and the
// per-axis min / max of the two-plane form are not needed.
__device__ __forceinline__ float box_near_oct(float nx, float fx, float ny, float fy, float nz, float fz, V3 on, V3 of,
                                              V3 id, float tmin, float tmax, bool& hit) {
    float nr = fmaxf(fmaxf(fmaf(nx, id.x, -on.x), fmaf(ny, id.y, -on.y)), fmaxf(fmaf(nz, id.z, -on.z), tmin));
    float fr = fminf(fminf(fmaf(fx, id.x, -of.x), fmaf(fy, id.y, -of.y)), fminf(fmaf(fz, id.z, -of.z), tmax));
    hit = nr <= fr * CRT_BOX_FAR;
    return nr;
}

__device__ __forceinline__ void cswap(float& ka, int32_t& va, float& kb, int32_t& vb) {
    bool s = kb < ka;
    float k = s ? kb : ka; kb = s ? ka : kb; ka = k;
    int32_t v = s ? vb : va; vb = s ? va : vb; va = v;
}

#define CRT_EMPTY_CHILD 0x7fffffff

// Closest hit over the BVH (WIDTH 2: both child boxes in a 64-B parent node; WIDTH 4: four child
// boxes in a 128-B SoA node, children visited near-first). TWO_LEVEL enables instance leaves that
// move the ray to object space and descend into the instance's BLAS. The traversal is a resumable
// state machine: step() performs one node or leaf visit, so a persistent kernel can interleave
// traversal steps with fetching new rays for finished lanes.
template <int BLOCK, int RS, int LS, int GS, bool COUNT, int WIDTH, int TL, bool EXACT> struct Traverser {
    // TL: 0 one level (analytic shapes among the quads), 1 two levels without shapes, 2 two levels
    // with shapes (the shape test costs the two-level loop registers, so it is compiled in only
    // where a scene of the launch has shapes)
    static constexpr bool TWO_LEVEL = TL != 0;
    static constexpr bool SHAPES = TL != 1;
    Stack<BLOCK, RS, LS, GS> st;
    V3 o, d, id, ow, dw;
    uint32_t oct;          // bit a set where inv_d[a] < 0: the hi plane of axis a is the near one;
                           // bit 3: tp is set up for the current ray
    TriPre tp;
    HitRec best;
    uint32_t cur_inst;
    bool in_blas, camera;
    int blas_level;
    int32_t node;
    float tmin;

    // hardware reciprocal (1 ulp; the box test's far factor covers it): one instruction instead of
    // the ten of a correctly rounded division, paid again at every instance entry and exit
    static __device__ __forceinline__ float inv_of(float x) {
        return __builtin_amdgcn_rcpf(fabsf(x) > 1e-30f ? x : copysignf(1e-30f, x));
    }
    __device__ __forceinline__ void set_ray(V3 ro, V3 rd) {
        o = ro; d = rd;
        id = v3(inv_of(d.x), inv_of(d.y), inv_of(d.z));
        oct = (id.x < 0.0f ? 1u : 0u) | (id.y < 0.0f ? 2u : 0u) | (id.z < 0.0f ? 4u : 0u);
        // two-level traversal: the triangle test's shear (three divisions) is set up at the first
        // triangle leaf after a ray change, since most instance entries and exits never test a
        // triangle (3.7 % faster on the instanced scene; one-level kernels lose 1-2 % to the test)
        if (TWO_LEVEL) return;
        tp = tri_pre(d);
        oct |= 8u;
    }
    __device__ __forceinline__ void init(const SceneView& S, V3 ro, V3 rd, float t0, float t1, bool cam,
                                         uint32_t* lds_col, GSlot gstk) {
        st.init(lds_col, gstk);
        ow = ro; dw = rd;
        set_ray(ro, rd);
        best.t = t1; best.prim = CRT_MISS; best.inst = CRT_NO_INST; best.u = 0.0f; best.v = 0.0f; best.key = 0u;
        cur_inst = CRT_NO_INST;
        in_blas = false; camera = cam;
        blas_level = 0;
        node = (int32_t)S.root;
        tmin = t0;
    }
    // Four-wide node on its near / far planes (picked by the ray's octant) and child refs: box
    // tests, near-first order, pushes; sets `node` to the nearest hit child (next = false) or
    // leaves next set when no child is hit. Empty child slots have boxes at +inf, which no ray hits.
    __device__ __forceinline__ void visit4(float4 rff, float4 nx, float4 fx, float4 ny, float4 fy, float4 nz, float4 fz,
                                           bool& next) {
        const float tf = best.t * CRT_TIE_HI;
        const int4 rf = make_int4((int)fbits(rff.x), (int)fbits(rff.y), (int)fbits(rff.z), (int)fbits(rff.w));
        bool h0, h1, h2, h3;
        // the plane offsets are formed at every visit: kept for the whole ray (hoisted out of the
        // loop) they take six registers, and the one-level kernel then spills inside its loop (5 %
        // slower than the two-operation planes); the empty asm keeps the compiler from hoisting
        float ox = o.x, oy = o.y, oz = o.z;
        asm volatile("" : "+v"(ox), "+v"(oy), "+v"(oz));
        const V3 oid = v3(ox * id.x, oy * id.y, oz * id.z);
        const V3 e = v3(fabsf(oid.x) * 2.384185791e-7f, fabsf(oid.y) * 2.384185791e-7f, fabsf(oid.z) * 2.384185791e-7f);
        const V3 on = oid + e, of = oid - e;
        float k0 = box_near_oct(nx.x, fx.x, ny.x, fy.x, nz.x, fz.x, on, of, id, tmin, tf, h0);
        float k1 = box_near_oct(nx.y, fx.y, ny.y, fy.y, nz.y, fz.y, on, of, id, tmin, tf, h1);
        float k2 = box_near_oct(nx.z, fx.z, ny.z, fy.z, nz.z, fz.z, on, of, id, tmin, tf, h2);
        float k3 = box_near_oct(nx.w, fx.w, ny.w, fy.w, nz.w, fz.w, on, of, id, tmin, tf, h3);
        const float INF = __builtin_inff();
        k0 = h0 ? k0 : INF;
        k1 = h1 ? k1 : INF;
        k2 = h2 ? k2 : INF;
        k3 = h3 ? k3 : INF;
        int32_t c0 = rf.x, c1 = rf.y, c2 = rf.z, c3 = rf.w;
        // near-first: sort the four (distance, child) pairs (missed children sort last at +inf),
        // push the hit ones far to near, continue with the nearest
        cswap(k0, c0, k1, c1); cswap(k2, c2, k3, c3);
        cswap(k0, c0, k2, c2); cswap(k1, c1, k3, c3);
        cswap(k1, c1, k2, c2);
        if (k0 < INF) {
            if (k3 < INF) st.push((uint32_t)c3);
            if (k2 < INF) st.push((uint32_t)c2);
            if (k1 < INF) st.push((uint32_t)c1);
            node = c0; next = false;
        }
    }
    // Leaf (reference written inline in the parent: no leaf-record fetch). With PRE, the first
    // triangle's record (v0, v1, v2 in a0-a2), the first sphere (a0) or the instance record
    // (a0-a3) is already loaded.
    template <bool PRE>
    __device__ __forceinline__ void visit_leaf(const SceneView& S, TravCount& tc, bool& next, float4 a0, float4 a1,
                                               float4 a2, float4 a3) {
        const uint32_t type = CRT_LEAF_REF_TYPE(node);
        const uint32_t cnt = CRT_LEAF_REF_COUNT(node);
        const uint32_t first = CRT_LEAF_REF_FIRST(node);
        if (COUNT) tc.leaves++;
        if (type == CRT_PRIM_TRI) {
            if (TWO_LEVEL && !(oct & 8u)) { tp = tri_pre(d); oct |= 8u; }
            for (uint32_t k = 0; k < cnt; ++k) {
                const uint32_t to = S.tri_off + (first + k) * 48u;
                float t, bu, bv;
                if (COUNT) tc.tris++;
                const bool pre = PRE && k == 0u;
                const float4 va = pre ? a0 : ldg_at(S.nodes, to);
                bool amb = false;
                if (tri_hit<EXACT>(o, d, tp, va, pre ? a1 : ldg_at(S.nodes, to + 16u),
                                   pre ? a2 : ldg_at(S.nodes, to + 32u), tmin, best.t * CRT_TIE_HI, t, bu, bv, amb) &&
                    hit_wins(t, fbits(va.w), best)) {
                    best.t = t; best.u = bu; best.v = bv; best.prim = (CRT_PRIM_TRI << 30) | (first + k);
                    best.inst = cur_inst; best.key = fbits(va.w);
                }
                // an undecided test (not EXACT) ends the query: a negative distance culls every box
                // and primitive left, and marks the result for the exact re-trace
                if (!EXACT && amb) best.t = -1.0f;
            }
        } else if (type == CRT_PRIM_SPHERE) {
            for (uint32_t k = 0; k < cnt; ++k) {
                const bool pre = PRE && k == 0u;
                const uint32_t so = S.sph_off + (first + k) * 32u;
                const float4 s = pre ? a0 : ldg_at(S.nodes, so);
                const float4 s1 = pre ? a1 : ldg_at(S.nodes, so + 16u);
                const uint32_t key = fbits(s1.x);
                float t;
                if (COUNT) tc.spheres++;
                // sphere lights hide from camera rays when invisible (omni_light.cuh:62-65)
                if ((fbits(s1.z) & CRT_SF_INVISIBLE) && camera) continue;
                if (sphere_hit(o, d, s, tmin, best.t * CRT_TIE_HI, t) && hit_wins(t, key, best)) {
                    best.t = t; best.u = 0.0f; best.v = 0.0f; best.prim = (CRT_PRIM_SPHERE << 30) | (first + k);
                    best.inst = cur_inst; best.key = key;
                }
            }
        } else if (type == CRT_PRIM_QUAD) {
            for (uint32_t k = 0; k < cnt; ++k) {
                float t, qa, qb;
                uint32_t key;
                if (COUNT) tc.quads++;
                if (quad_hit<SHAPES>(o, d, S.quad(first + k), tmin, best.t * CRT_TIE_HI, camera, t, qa, qb, key) &&
                    hit_wins(t, key, best)) {
                    best.t = t; best.u = qa; best.v = qb; best.prim = (CRT_PRIM_QUAD << 30) | (first + k);
                    best.inst = cur_inst; best.key = key;
                }
            }
        } else if (TWO_LEVEL && !in_blas) {
            // instance leaf (count 1): move the ray to object space and descend into the BLAS
            const uint32_t io = S.inst_off + first * 64u;
            const float4 r0 = PRE ? a0 : ldg_at(S.nodes, io), r1 = PRE ? a1 : ldg_at(S.nodes, io + 16u);
            const float4 r2 = PRE ? a2 : ldg_at(S.nodes, io + 32u), m4 = PRE ? a3 : ldg_at(S.nodes, io + 48u);
            const int4 meta = make_int4((int)fbits(m4.x), (int)fbits(m4.y), (int)fbits(m4.z), (int)fbits(m4.w));
            if (COUNT) tc.insts++;
            (void)meta.y;
            set_ray(v3(r0.x * ow.x + r0.y * ow.y + r0.z * ow.z + r0.w, r1.x * ow.x + r1.y * ow.y + r1.z * ow.z + r1.w,
                       r2.x * ow.x + r2.y * ow.y + r2.z * ow.z + r2.w),
                    v3(r0.x * dw.x + r0.y * dw.y + r0.z * dw.z, r1.x * dw.x + r1.y * dw.y + r1.z * dw.z,
                       r2.x * dw.x + r2.y * dw.y + r2.z * dw.z));
            cur_inst = first;   // instance record index (leaf order); the host maps it to the caller's id
            in_blas = true;
            blas_level = st.depth();
            node = meta.x;
            next = false;
        }
    }
    // One visit; returns true when the traversal is complete (result in `best`).
    __device__ __forceinline__ bool step(const SceneView& S, TravCount& tc) {
        bool next = true;
        if (WIDTH == 4) {
            // Node lanes and leaf lanes of a wave share one set of seven loads (one memory round
            // trip per step instead of one per branch taken): node lanes fetch their near / far
            // planes and child refs, triangle lanes their (first) triangle record, sphere lanes
            // their (first) sphere record (both halves) and instance lanes their instance record, all addressed from
            // the node array. Quad lanes read node 0 and load their quads in the leaf branch.
            const bool isn = node >= 0;
            const uint32_t ltype = CRT_LEAF_REF_TYPE(node), lfirst = CRT_LEAF_REF_FIRST(node);
            const bool ist = !isn && ltype == CRT_PRIM_TRI;
            const bool iss = !isn && ltype == CRT_PRIM_SPHERE;
            const bool isi = TWO_LEVEL && !isn && ltype == CRT_PRIM_INST && !in_blas;
            const uint32_t lb = ist ? S.tri_off + lfirst * 48u
                                    : (iss ? S.sph_off + lfirst * 32u : (isi ? S.inst_off + lfirst * 64u : 0u));
            const uint32_t l3 = isi ? lb + 48u : lb;
            const uint32_t nb = (uint32_t)node * 128u;   // node byte offset (scenes stay below 4 GiB of geometry)
            const uint32_t sx = (oct & 1u) << 4, sy = (oct & 2u) << 3, sz = (oct & 4u) << 2;
            float4 rff = ldg_at(S.nodes, isn ? nb + 96u : lb);
            float4 nx = ldg_at(S.nodes, isn ? nb + sx : lb + 16u), fx = ldg_at(S.nodes, isn ? nb + (16u - sx) : lb + 32u),
                   ny = ldg_at(S.nodes, isn ? nb + 32u + sy : l3), fy = ldg_at(S.nodes, isn ? nb + (48u - sy) : lb),
                   nz = ldg_at(S.nodes, isn ? nb + 64u + sz : lb), fz = ldg_at(S.nodes, isn ? nb + (80u - sz) : lb);
            __builtin_amdgcn_sched_group_barrier(0x020, 7, 0);   // 0x020: VMEM reads
            if (isn) {
                if (COUNT) tc.nodes++;
                visit4(rff, nx, fx, ny, fy, nz, fz, next);
            } else {
                visit_leaf<true>(S, tc, next, rff, nx, fx, ny);
            }
        } else if (node >= 0) {
            if (COUNT) tc.nodes++;
            if (WIDTH == 2) {
                const float tf = best.t * CRT_TIE_HI;
                const float4* np = S.nodes + 4 * (uint32_t)node;
                float4 n0 = ldg(np), n1 = ldg(np + 1), n2 = ldg(np + 2);
                int4 n3 = ldg_i4(np + 3);
                bool hit0, hit1;
                float near0 = box_near(n0.x, n0.y, n0.z, n0.w, n2.x, n2.y, o, id, tmin, tf, hit0);
                float near1 = box_near(n1.x, n1.y, n1.z, n1.w, n2.z, n2.w, o, id, tmin, tf, hit1);
                if (hit0 && hit1) {
                    bool swap = near1 < near0;
                    int32_t a = swap ? n3.y : n3.x, b = swap ? n3.x : n3.y;
                    st.push((uint32_t)b);
                    node = a; next = false;
                } else if (hit0) { node = n3.x; next = false; }
                else if (hit1) { node = n3.y; next = false; }
            } else {
                // planes lo.x hi.x lo.y hi.y lo.z hi.z (four children each) and the child refs; the
                // near and far plane of each axis are picked by the ray's octant (no per-axis min /
                // max); the seven loads are kept together ahead of the box tests (one memory round
                // trip per visit: the scheduler would otherwise place the ref load late to save
                // registers)
                const uint32_t nb = (uint32_t)node * 128u;   // node byte offset (scenes stay below 4 GiB of nodes)
                const uint32_t sx = (oct & 1u) << 4, sy = (oct & 2u) << 3, sz = (oct & 4u) << 2;
                float4 rff = ldg_at(S.nodes, nb + 96u);
                float4 nx = ldg_at(S.nodes, nb + sx), fx = ldg_at(S.nodes, nb + (16u - sx)),
                       ny = ldg_at(S.nodes, nb + 32u + sy), fy = ldg_at(S.nodes, nb + (48u - sy)),
                       nz = ldg_at(S.nodes, nb + 64u + sz), fz = ldg_at(S.nodes, nb + (80u - sz));
                __builtin_amdgcn_sched_group_barrier(0x020, 7, 0);   // 0x020: VMEM reads
                visit4(rff, nx, fx, ny, fy, nz, fz, next);
            }
        } else {
            const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            visit_leaf<false>(S, tc, next, z, z, z, z);
        }
        if (next) {
            if (TWO_LEVEL && in_blas && st.depth() == blas_level) {
                set_ray(ow, dw);
                in_blas = false; cur_inst = CRT_NO_INST;
            }
            if (st.empty()) return true;
            node = (int32_t)st.pop();
        }
        return false;
    }
};

template <int BLOCK, int RS, int LS, int GS, bool COUNT, int WIDTH, int TL, bool EXACT>
__device__ __forceinline__ void trace(const SceneView& S, V3 ow, V3 dw, float tmin, float tmax, bool camera,
                                      uint32_t* lds_col, GSlot gstk, HitRec& best, TravCount& tc, bool& ovf) {
    Traverser<BLOCK, RS, LS, GS, COUNT, WIDTH, TL, EXACT> tr;
    tr.init(S, ow, dw, tmin, tmax, camera, lds_col, gstk);
    while (!tr.step(S, tc)) {
    }
    best = tr.best;
    ovf = tr.st.overflow;
}

// ------------------------------------------------------------------ block-wide append (compaction)
// Each wave ballots its keepers, lane offsets come from mbcnt, the block's wave totals are scanned
// in LDS and one atomicAdd per block reserves the output range.
template <int BLOCK> __device__ __forceinline__ uint32_t block_append(bool keep, uint32_t* counter, uint32_t* s_wave,
                                                                    uint32_t* s_base) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int NW = BLOCK / 64;
    uint64_t m = __ballot(keep);
    uint32_t lo = (uint32_t)m, hi = (uint32_t)(m >> 32);
    uint32_t rank = __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
    if (lane == 0) s_wave[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int w = 0; w < NW; ++w) { uint32_t c = s_wave[w]; s_wave[w] = run; run += c; }
        *s_base = run ? atomicAdd(counter, run) : 0u;
    }
    __syncthreads();
    uint32_t pos = *s_base + s_wave[wave] + rank;
    __syncthreads();
    return pos;
}

// Block-wide append with the keepers grouped by a 3-bit key (the direction octant of the next
// segment): within the block's output range, keepers are ordered key-major, then wave, then lane.
// Waves of the next closest-hit launch then hold rays of neighbouring paths that share an octant,
// so their near-first traversal orders agree. Deterministic (ballot + mbcnt ranks, one atomic per
// block). s_cnt holds 8 * BLOCK/64 words.
template <int BLOCK> __device__ __forceinline__ uint32_t block_append_sorted(bool keep, uint32_t key, uint32_t* counter,
                                                                           uint32_t* s_cnt, uint32_t* s_base) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int NW = BLOCK / 64;
    uint32_t rank = 0;
#pragma unroll
    for (uint32_t o = 0; o < 8; ++o) {
        uint64_t m = __ballot(keep && key == o);
        if (key == o) rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (lane == 0) s_cnt[o * NW + wave] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int e = 0; e < 8 * NW; ++e) { uint32_t c = s_cnt[e]; s_cnt[e] = run; run += c; }
        *s_base = run ? atomicAdd(counter, run) : 0u;
    }
    __syncthreads();
    uint32_t pos = *s_base + s_cnt[(key & 7u) * NW + wave] + rank;
    __syncthreads();
    return pos;
}

// Adds every lane's v to *p with one atomic per wave (a per-lane atomic on one address serialises
// in L2). All lanes of the wave must call it.
__device__ __forceinline__ void wave_add(unsigned long long* p, uint32_t v) {
    uint32_t s = v;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += (uint32_t)__shfl_xor((int)s, off);
    if ((threadIdx.x & 63u) == 0u && s) atomicAdd(p, (unsigned long long)s);
}

}  // namespace

// ======================================================================================== kernels

// Path state (SoA over path slots):
//   ray_o  float4 (origin, time)           ray_d float4 (dir, state: depth | bounce << 8 | camera << 16
//                                                                 | pass-through re-traces so far << 24)
//   thr    float4 (throughput, minstd state) rad   float4 (radiance sum, -)
//   ids    uint2  (storage pixel, request << 16 | sample-in-request)
//   hit    float4 (t, prim bits, u, v)      hit_inst uint32

#define ST_DEPTH(s) ((s)&0xffu)
#define ST_BOUNCE(s) (((s) >> 8) & 0xffu)
#define ST_CAMERA 0x10000u
#define ST_RETRACES(s) ((s) >> 24)

__device__ __forceinline__ int find_request(const crt_dev_request* reqs, int nreq, uint32_t p) {
    int lo = 0, hi = nreq - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (reqs[mid].path_base <= p) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Camera ray of (pixel, sample): render kernel sample loop (core/src/gpu.cu:776-787) +
// perspective_camera::get_ray (core/src/cameras/perspective_camera.cuh:106-113): jitter r,
// diagonal stratified (u, v), lens disk by rejection, time draw. `g` is positioned at the start
// of the sample's draws.
template <int MODE>
__device__ __forceinline__ void camera_ray(const crt_dev_request& R, uint32_t pix, uint32_t sample, Rng<MODE>& g,
                                           V3& o, V3& dir, float& tm) {
    uint32_t W = R.width, H = R.height;
    uint32_t jj = pix / W, i = pix - jj * W;
    uint32_t j = H - 1u - jj;                  // the reference's row index counts from the bottom
    int sq = (int)R.sqrt_spp;
    int sj = (int)(sample / (uint32_t)sq), si = (int)(sample - (uint32_t)sj * (uint32_t)sq);
    float rr = g.uniform();
    float u = ((float)i + ((float)si + rr) / (float)sq) / (float)W;
    float v = ((float)j + ((float)sj + rr) / (float)sq) / (float)H;
    float dx, dy;
    do {
        float a = g.uniform();
        float b = g.uniform();
        dx = 2.0f * a - 1.0f; dy = 2.0f * b - 1.0f;
    } while (dx * dx + dy * dy + 0.0f * 0.0f >= 1.0f);
    V3 rd = R.lens_radius * v3(dx, dy, 0.0f);
    V3 cu = v3p(R.cu), cv = v3p(R.cv), org = v3p(R.origin);
    V3 off = cu * rd.x + cv * rd.y;
    tm = R.time0 + g.uniform() * (R.time1 - R.time0);
    if (R.ortho) {
        // orthographic_camera::get_ray (core/src/cameras/orthographic_camera.cuh:125-142): the
        // viewport point plus the lens offset, along -w
        o = v3p(R.llc) + u * v3p(R.horizontal) + v * v3p(R.vertical) + off;
        dir = v3p(R.ortho_dir);
        return;
    }
    o = org + off;
    dir = v3p(R.llc) + u * v3p(R.horizontal) + v * v3p(R.vertical) - org - off;
}

// Pixel index in the reference's numbering (rows counted from the bottom), which keys the streams.
__device__ __forceinline__ uint32_t ref_pixel_of(const crt_dev_request& R, uint32_t pix) {
    uint32_t jj = pix / R.width, i = pix - jj * R.width;
    return (R.height - 1u - jj) * R.width + i;
}

// Streaming pass control block (CRT_CTL_* words, crt_kernels.h): two halves alternate between
// iterations. Iteration k reads half k%2 ("in": survivors left by the previous shade stage, next
// path id to generate) and prepares half (k+1)%2 ("out") for its own shade stage. Live paths are
// kept compact: the shade stage writes survivors to the front of the other state buffer and the
// next regeneration appends new camera paths behind them, so every state access is coalesced.
__global__ void k_pass_init(uint32_t* __restrict__ ctl, uint32_t nwords) {
    for (uint32_t i = threadIdx.x; i < nwords; i += blockDim.x) ctl[i] = 0u;
}

// Image pixel of the s-th pixel slot of a work item: when the item covers the whole image and both
// sides are multiples of 8, 8 x 8 tiles in row-major tile order (64 consecutive slots trace one
// tile instead of a 64-pixel row strip: more coherent camera rays); otherwise row-major. The path
// carries the image pixel from here on (RNG key, radiance slot, accumulation), so the image does not
// depend on this order.
__device__ __forceinline__ uint32_t item_pixel(const crt_dev_request& R, uint32_t s) {
    const uint32_t W = R.width, H = R.height;
    if (R.pix_begin != 0u || R.pix_count != W * H || ((W | H) & 7u) != 0u) return R.pix_begin + s;
    const uint32_t tile = s >> 6, l = s & 63u, tpr = W >> 3;
    const uint32_t ty = tile / tpr, tx = tile - ty * tpr;
    return (ty * 8u + (l >> 3)) * W + tx * 8u + (l & 7u);
}

// Regeneration: the slots behind the survivors take the next path ids of the (sub-)pass, whose
// ids are [pid0, pid0 + total) (camera rays, in path-id order: pixel-major within each work item,
// the item's samples of one pixel consecutive). ONE: a closed pass of one request (its record is
// wave-uniform: scalar loads, no search).
template <int MODE, bool ONE>
__global__ __launch_bounds__(256) void k_regen(const crt_dev_request* __restrict__ reqs, int nreq, uint32_t pid0,
                                               uint32_t total, uint32_t nslots, uint32_t* __restrict__ ctl_in,
                                               uint32_t* __restrict__ ctl_out, const uint32_t* __restrict__ open_pass,
                                               crt_paths P) {
    // an open pass (serving) grows while it runs: its path count and request count live in device
    // memory, written by the host between iterations
    if (open_pass) { total = open_pass[0]; nreq = (int)open_pass[1]; }
    const uint32_t surv = ctl_in[CRT_CTL_SURV], g0 = ctl_in[CRT_CTL_GEN];
    const uint32_t n_new = min(nslots - surv, total - g0);
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid == 0) {
        ctl_in[CRT_CTL_LIVE] = surv + n_new;
        ctl_in[CRT_CTL_WORK] = 0u;
        ctl_out[CRT_CTL_GEN] = g0 + n_new;
        ctl_out[CRT_CTL_SURV] = 0u;
    }
    for (uint32_t i = tid; i < n_new; i += gridDim.x * blockDim.x) {
        const uint32_t slot = surv + i, pid = pid0 + g0 + i;
        const int r = ONE ? 0 : find_request(reqs, nreq, pid);
        const crt_dev_request& R = reqs[r];
        uint32_t local = pid - R.path_base;
        // pixel-major: the samples of a pixel are consecutive path ids (a wave traces 64 samples of
        // one pixel at high spp: nearly identical camera rays)
        const uint32_t sidx = local / R.sample_count, k = local - sidx * R.sample_count;
        uint32_t pix = item_pixel(R, sidx);
        uint32_t sample = R.sample_begin + k;
        Rng<MODE> g;
        g.dyn = R.rng_mode;
        g.ms = 0u;
        if (g.minstd()) g.ms = ldg(R.rng_state + pix);
        else g.pcg = pcg_keyed(R.seed, ref_pixel_of(R, pix), sample, 0u);
        V3 o, dir;
        float tm;
        camera_ray<MODE>(R, pix, sample, g, o, dir, tm);
        uint32_t state = (R.max_depth & 0xffu) | ST_CAMERA;
        P.ray_o[slot] = make_float4(o.x, o.y, o.z, tm);
        P.ray_d[slot] = make_float4(dir.x, dir.y, dir.z, bitsf(state));
        P.thr[slot] = make_float4(1.0f, 1.0f, 1.0f, bitsf(g.ms));
        P.rad[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        P.ids[slot] = make_uint2(pix, ((uint32_t)r << 16) | (sample & 0xffffu));
    }
}

// Closest hit of one path segment. One-sided emitters hit from behind are re-traced from
// t + 0.001 (core/src/cameras/camera.cuh:179-183); the result is then flagged as a pass-through
// (returned flag CRT_PASSTHROUGH). `nrays` counts closest-hit queries, `nretrace` the re-traces.
// A query left undecided by a non-EXACT triangle test (its result distance set to -1) returns with
// `amb` set and counts nothing: the caller re-traces it with an EXACT traversal, which answers and
// counts it.
template <int BLOCK, int RS, int LS, int GS, bool COUNT, int WIDTH, int TL, bool EXACT>
__device__ __forceinline__ uint32_t closest_hit(const SceneView& S, V3 o, V3 d, bool camera, uint32_t* lds_col,
                                                GSlot gst, HitRec& h, TravCount& tc, bool& ovf, uint32_t& nrays,
                                                uint32_t& nretrace, bool& amb) {
    float tmin = 0.00001f;
    uint32_t pass_flag = 0u;
    amb = false;
    for (;;) {
        bool o1 = false;
        trace<BLOCK, RS, LS, GS, COUNT, WIDTH, TL, EXACT>(S, o, d, tmin, 3.402823466e+38f, camera, lds_col, gst, h, tc,
                                                                 o1);
        ovf = ovf || o1;
        if (!EXACT && h.t < 0.0f) { amb = true; return 0u; }
        if (pass_flag) break;
        const uint32_t pf = passes_through(S, h, o, d);
        if (!pf) break;
        tmin = h.t + 0.001f;
        pass_flag = pf;
    }
    if (pass_flag) nretrace++;
    nrays++;
    return pass_flag;
}

// The exact re-trace of a query the float triangle test left undecided, by the lane that met it
// after its float traversal (rare and divergent; the double-precision code stays out of the
// traversal loop). Out of line it would be allocated without the kernel's register bound: the
// whole kernel would then take its count (180-210 VGPRs, two waves per SIMD).
struct ExactHit { HitRec h; uint32_t pass_flag, ovf; };
template <int BLOCK, int RS, int LS, int GS, int WIDTH, int TL>
__device__ __forceinline__ ExactHit exact_hit(const SceneView& S, V3 o, V3 d, bool camera, uint32_t* lds_col, GSlot gst) {
    ExactHit r;
    TravCount tc = {0, 0, 0, 0, 0, 0};
    bool ovf = false, amb = false;
    uint32_t nrays = 0, nretrace = 0;
    r.pass_flag = closest_hit<BLOCK, RS, LS, GS, false, WIDTH, TL, true>(S, o, d, camera, lds_col, gst, r.h, tc, ovf,
                                                                               nrays, nretrace, amb);
    r.ovf = ovf ? 1u : 0u;
    return r;
}

// Closest hit with the float triangle test and, for the rare query it leaves undecided, the exact
// re-trace (counted in nexact); answers and counts as closest_hit<EXACT = true> would.
template <int BLOCK, int RS, int LS, int GS, bool COUNT, int WIDTH, int TL>
__device__ __forceinline__ uint32_t closest_hit_checked(const SceneView& S, V3 o, V3 d, bool camera, uint32_t* lds_col,
                                                        GSlot gst, HitRec& h, TravCount& tc, uint32_t& novf,
                                                        uint32_t& nrays, uint32_t& nretrace, uint32_t& nexact) {
    bool ovf = false, amb = false;
    uint32_t pass_flag = closest_hit<BLOCK, RS, LS, GS, COUNT, WIDTH, TL, COUNT>(S, o, d, camera, lds_col, gst, h, tc,
                                                                                      ovf, nrays, nretrace, amb);
    novf += ovf ? 1u : 0u;
    if (!COUNT && amb) {
        const ExactHit e = exact_hit<BLOCK, RS, LS, GS, WIDTH, TL>(S, o, d, camera, lds_col, gst);
        h = e.h;
        pass_flag = e.pass_flag;
        novf += e.ovf;
        nrays++;
        if (pass_flag) nretrace++;
        ++nexact;
    }
    return pass_flag;
}

// Closest-hit stage over the live path queue.
// UNIFORM: every ray of the launch belongs to scene `uniform_scene` (scene pointers stay scalar).
// Minimum resident blocks per CU the register allocation must allow. Blocks are 256 threads, one
// wave on each of a CU's 4 SIMDs, so this is also waves per SIMD. The traversal is latency-bound and
// occupancy pays more than spills outside the traversal loop, not more than spills inside it
// (measured): the single-scene one-level kernel is held to 6 (80 VGPRs, no spill in the loop); the
// multi-scene one-level kernel, whose scene offsets are per-lane values, to 5 (96 VGPRs: at 6 the
// loop spills, and a 64-scene mixed batch and the saturated serving mix run 9-10 % slower);
// two-level kernels to 5 (96 VGPRs: 6 % faster than at 4; at 6 they spill inside the loop and lose
// 17 %). Counting variants are not bounded.
constexpr int extend_min_waves(bool count, bool uniform, bool two_level) {
    return count ? 1 : (two_level ? 5 : (uniform ? 6 : 5));
}

template <int BLOCK, int RS, int LS, int GS, bool COUNT, bool UNIFORM, int WIDTH, int TL>
__global__ __launch_bounds__(BLOCK, extend_min_waves(COUNT, UNIFORM, TL != 0)) void k_extend(const crt_dev_scene* __restrict__ scenes,
                                                  const crt_dev_request* __restrict__ reqs, uint32_t uniform_scene,
                                                  const uint32_t* __restrict__ queue, const uint32_t* __restrict__ count_ptr,
                                                  uint32_t count_direct, const float4* __restrict__ ray_o,
                                                  const float4* __restrict__ ray_d, const uint2* __restrict__ ids,
                                                  float4* __restrict__ hit, uint32_t* __restrict__ hit_inst,
                                                  uint32_t* __restrict__ gstack, uint32_t* __restrict__ zero_counter,
                                                  uint32_t* __restrict__ work_counter, uint32_t chunk_groups,
                                                  crt_dev_counters* __restrict__ counters) {
    __shared__ uint32_t s_stack[LS * BLOCK];
    if (zero_counter && blockIdx.x == 0 && threadIdx.x == 0) *zero_counter = 0u;
    const uint32_t n = count_ptr ? *count_ptr : count_direct;
    const uint32_t gtid = blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t stride = gridDim.x * BLOCK;
    uint32_t* lds_col = s_stack + threadIdx.x;
    const GSlot gst{gstack, gtid * GS};
    TravCount tc = {0, 0, 0, 0, 0, 0};
    uint32_t nrays = 0, nretrace = 0, novf = 0, nexact = 0;
    SceneView SU;
    if (UNIFORM) SU = scene_view(scenes + uniform_scene);
    auto one_ray = [&](uint32_t i) {
        uint32_t slot = queue ? queue[i] : i;
        float4 ro = ray_o[slot], rdv = ray_d[slot];
        uint32_t state = fbits(rdv.w);
        SceneView S;
        if (UNIFORM) S = SU;
        else S = scene_view(scenes + reqs[ids[slot].y >> 16].scene);
        V3 o = v3(ro), d = v3(rdv);
        bool camera = (state & ST_CAMERA) != 0;
        HitRec h;
        const uint32_t pass_flag = closest_hit_checked<BLOCK, RS, LS, GS, COUNT, WIDTH, TL>(
            S, o, d, camera, lds_col, gst, h, tc, novf, nrays, nretrace, nexact);
        hit[slot] = make_float4(h.t, bitsf(h.prim), h.u, h.v);
        hit_inst[slot] = h.inst | pass_flag;
    };
    if (work_counter) {
        // persistent waves (grid sized to the resident waves): each wave takes the next chunk x 64
        // consecutive rays when all its lanes are done, so no wave idles while another still holds
        // a backlog; the fetch and the exit test are wave-uniform
        const uint32_t lane = threadIdx.x & 63u;
        for (;;) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(work_counter, 64u * chunk_groups);
            base = (uint32_t)__shfl((int)base, 0);
            if (base >= n) break;
            const uint32_t wend = min(base + 64u * chunk_groups, n);
            for (uint32_t i = base + lane; i < wend; i += 64u) one_ray(i);
        }
    } else {
        for (uint32_t i = gtid; i < n; i += stride) one_ray(i);
    }
    wave_add(&counters->rays, nrays);
    wave_add(&counters->retraces, nretrace);
    if (COUNT) {
        wave_add(&counters->node_visits, tc.nodes);
        wave_add(&counters->leaf_visits, tc.leaves);
        wave_add(&counters->tri_tests, tc.tris);
        wave_add(&counters->sphere_tests, tc.spheres);
        wave_add(&counters->quad_tests, tc.quads);
        wave_add(&counters->inst_visits, tc.insts);
    }
    wave_add(&counters->stack_overflow, novf);
    wave_add(&counters->exact, nexact);
}

// Closest-hit stage with per-lane refill for the survivors of the last shade stage: every wave owns
// a contiguous share of them, and a lane whose ray is finished takes the wave's next ray instead of
// waiting for the slowest lane of the wave. A refill is made when at least `refill_min` lanes are
// idle (or all are): the idle lanes take the next rays of the share in lane order (mbcnt rank), so
// refilled rays are neighbours of each other and of the rays in flight. Camera rays (behind the
// survivors when `first_camera` is given) keep whole-wave chunks in grid-stride order, whose
// coherence refills would break. Share and chunk pointers are wave-uniform and private: no atomics.
// A wave leaves when its share and chunks are handed out and its lanes are idle; each traversal
// ends (every node is visited at most once per query, at most one re-trace). Pays when rays of one
// wave differ much in length (many primitive tests per ray, little coherence).
// one-level: 6 (80 VGPRs, a few per-ray spills; 2.5 % faster than at 5); two-level: 5 (96 VGPRs,
// 7 % faster on the instanced scene than unbounded at 4; at 6 the loop spills and it takes 1.6x
// as long)
constexpr int extend_pt_min_waves(bool count, bool uniform, bool two_level, int width) {
    return count ? 1 : (two_level ? 5 : 6);
}

template <int BLOCK, int RS, int LS, int GS, bool COUNT, bool UNIFORM, int WIDTH, int TL>
__global__ __launch_bounds__(BLOCK, extend_pt_min_waves(COUNT, UNIFORM, TL != 0, WIDTH)) void k_extend_pt(const crt_dev_scene* __restrict__ scenes,
                                                     const crt_dev_request* __restrict__ reqs, uint32_t uniform_scene,
                                                     const uint32_t* __restrict__ queue,
                                                     const uint32_t* __restrict__ count_ptr, uint32_t count_direct,
                                                     const float4* __restrict__ ray_o, const float4* __restrict__ ray_d,
                                                     const uint2* __restrict__ ids, float4* __restrict__ hit,
                                                     uint32_t* __restrict__ hit_inst, uint32_t* __restrict__ gstack,
                                                     uint32_t* __restrict__ zero_counter, uint32_t refill_min,
                                                     const uint32_t* __restrict__ first_camera,
                                                     crt_dev_counters* __restrict__ counters) {
    __shared__ uint32_t s_stack[LS * BLOCK];
    if (zero_counter && blockIdx.x == 0 && threadIdx.x == 0) *zero_counter = 0u;
    const uint32_t n = count_ptr ? *count_ptr : count_direct;
    const uint32_t cam0 = first_camera ? *first_camera : n;
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t* lds_col = s_stack + threadIdx.x;
    const GSlot gst{gstack, (blockIdx.x * BLOCK + threadIdx.x) * GS};
    TravCount tc = {0, 0, 0, 0, 0, 0};
    uint32_t nrays = 0, nretrace = 0, novf = 0, nexact = 0;
    SceneView S;
    if (UNIFORM) S = scene_view(scenes + uniform_scene);
    Traverser<BLOCK, RS, LS, GS, COUNT, WIDTH, TL, COUNT> tr;
    bool active = false, exhausted = false;
    uint32_t slot = 0, pass_flag = 0;
    // Survivors of the last shade stage, queue [0, nsec): this wave's contiguous share [nxt, wend),
    // idle lanes refilled as soon as refill_min of them are idle. Camera rays, [nsec, n): chunks of
    // 64 consecutive rays in grid-stride order (chunk c of the wave starts at nsec + (wid + c *
    // nwaves) * 64, so neighbouring waves trace neighbouring pixels, as in the grid-stride
    // kernel), a chunk taken when all 64 lanes are idle.
    const uint32_t nwaves = gridDim.x * (BLOCK / 64);
    const uint32_t wid = blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
    const uint32_t nsec = min(cam0, n);
    const uint32_t share = (nsec + nwaves - 1) / nwaves;
    uint32_t nxt = min(wid * share, nsec);
    const uint32_t wend = min(nxt + share, nsec);
    uint32_t cnext = nsec + wid * 64u;
    const uint32_t cstride = nwaves * 64u;
    bool camera_phase = nxt >= wend;
    for (;;) {
        if (!exhausted) {
            uint64_t idle = __ballot(!active);
            uint32_t cnt = (uint32_t)__popcll(idle);
            uint32_t i = 0xffffffffu, lim = 0u;
            if (camera_phase) {
                if (cnt == 64u) {
                    if (cnext >= n) {
                        exhausted = true;
                    } else {
                        i = cnext + lane;
                        lim = n;
                        cnext += cstride;
                    }
                }
            } else if (cnt >= refill_min || cnt == 64u) {
                const uint32_t base = nxt;
                nxt += cnt;
                if (nxt >= wend) camera_phase = true;
                i = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
                lim = wend;
            }
            if (!active && i < lim) {
                slot = queue ? queue[i] : i;
                float4 ro = ray_o[slot], rdv = ray_d[slot];
                if (!UNIFORM) S = scene_view(scenes + reqs[ids[slot].y >> 16].scene);
                tr.init(S, v3(ro), v3(rdv), 0.00001f, 3.402823466e+38f, (fbits(rdv.w) & ST_CAMERA) != 0, lds_col, gst);
                pass_flag = 0u;
                active = true;
            }
        }
        if (__ballot(active) == 0ull) {
            if (exhausted) break;
            continue;
        }
        if (active && tr.step(S, tc)) {
            const HitRec h = tr.best;
            novf += tr.st.overflow ? 1u : 0u;
            if (!COUNT && h.t < 0.0f) {
                // undecided triangle test: the lane re-traces the query exactly (rare, divergent)
                const ExactHit e = exact_hit<BLOCK, RS, LS, GS, WIDTH, TL>(S, tr.ow, tr.dw, tr.camera, lds_col, gst);
                novf += e.ovf;
                nrays++;
                if (e.pass_flag) nretrace++;
                ++nexact;
                hit[slot] = make_float4(e.h.t, bitsf(e.h.prim), e.h.u, e.h.v);
                hit_inst[slot] = e.h.inst | e.pass_flag;
                active = false;
                continue;
            }
            // an emitter seen from behind: re-trace from t + 0.001 (camera.cuh:179-183)
            const uint32_t again = pass_flag ? 0u : passes_through(S, h, tr.ow, tr.dw);
            if (again) {
                pass_flag = again;
                tr.init(S, tr.ow, tr.dw, h.t + 0.001f, 3.402823466e+38f, tr.camera, lds_col, gst);
            } else {
                nrays++;
                if (pass_flag) nretrace++;
                hit[slot] = make_float4(h.t, bitsf(h.prim), h.u, h.v);
                hit_inst[slot] = h.inst | pass_flag;
                active = false;
            }
        }
    }
    wave_add(&counters->rays, nrays);
    wave_add(&counters->retraces, nretrace);
    if (COUNT) {
        wave_add(&counters->node_visits, tc.nodes);
        wave_add(&counters->leaf_visits, tc.leaves);
        wave_add(&counters->tri_tests, tc.tris);
        wave_add(&counters->sphere_tests, tc.spheres);
        wave_add(&counters->quad_tests, tc.quads);
        wave_add(&counters->inst_visits, tc.insts);
    }
    wave_add(&counters->stack_overflow, novf);
    wave_add(&counters->exact, nexact);
}

// Closest hit of caller rays (first-hit parity queries).
template <int BLOCK, int RS, int LS, int GS, int WIDTH, int TL>
__global__ __launch_bounds__(BLOCK) void k_trace_rays(const crt_dev_scene* __restrict__ sc, const float4* __restrict__ rays,
                                                      uint32_t n, int camera, float4* __restrict__ hit,
                                                      uint32_t* __restrict__ hit_inst, uint32_t* __restrict__ gstack,
                                                      crt_dev_counters* __restrict__ counters) {
    __shared__ uint32_t s_stack[LS * BLOCK];
    const uint32_t gtid = blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t stride = gridDim.x * BLOCK;
    SceneView S = scene_view(sc);
    TravCount tc = {0, 0, 0, 0, 0, 0};
    uint32_t novf = 0;
    for (uint32_t i = gtid; i < n; i += stride) {
        float4 a = rays[2 * i], b = rays[2 * i + 1];
        HitRec h;
        bool ovf = false;
        trace<BLOCK, RS, LS, GS, false, WIDTH, TL, true>(S, v3(a), v3(b), a.w, b.w, camera != 0, s_stack + threadIdx.x,
                                                                GSlot{gstack, gtid * GS}, h, tc, ovf);
        novf += ovf ? 1u : 0u;
        hit[i] = make_float4(h.t, bitsf(h.prim), h.u, h.v);
        hit_inst[i] = h.inst;
    }
    wave_add(&counters->stack_overflow, novf);
}

// ------------------------------------------------------------------ shading
namespace {
struct Mat { float4 m0; int4 m1; };

__device__ __forceinline__ float4 tex_value(const float4* texs, const uint8_t* texels, int ti, float u, float v, V3 p) {
    for (int hop = 0; hop < 2; ++hop) {
        float4 t0 = ldg(texs + 2 * ti);
        int4 t1 = ldg_i4(texs + 2 * ti + 1);
        if (t1.x == CRT_TEX_SOLID) return t0;
        if (t1.x == CRT_TEX_CHECKER) {
            // checker_texture::value (core/src/textures/checker_texture.cuh:33-42); t0.w = 1/scale
            float inv = t0.w;
            int xi = (int)floorf(inv * p.x), yi = (int)floorf(inv * p.y), zi = (int)floorf(inv * p.z);
            bool even = ((xi + yi + zi) % 2) == 0;
            ti = even ? t1.y : t1.z;
            if (hop == 1) break;
            continue;
        }
        // image_texture::value (core/src/textures/image_texture.cuh:54-71)
        uint32_t W = (uint32_t)t1.y, H = (uint32_t)t1.z;
        uint32_t off = (uint32_t)fbits(t0.x), ch = (uint32_t)fbits(t0.y);
        if (H == 0) return make_float4(0.0f, 1.0f, 1.0f, 0.0f);
        float uu = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
        float vv = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
        vv = 1.0f - vv;
        int i = (int)(uu * (float)W), j = (int)(vv * (float)H);
        i = i < 0 ? 0 : (i < (int)W ? i : (int)W - 1);
        j = j < 0 ? 0 : (j < (int)H ? j : (int)H - 1);
        const uint8_t* px = texels + off + ((size_t)j * W + i) * ch;
        const float sc = 1.0f / 255.0f;
        if (ch == 4u) {
            uint32_t w4 = ldg((const uint32_t*)px);
            return make_float4(sc * (float)(w4 & 0xffu), sc * (float)((w4 >> 8) & 0xffu), sc * (float)((w4 >> 16) & 0xffu), 0.0f);
        }
        return make_float4(sc * (float)ldg(px), sc * (float)ldg(px + 1), sc * (float)ldg(px + 2), 0.0f);
    }
    return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}
}  // namespace

// Shading of one hit (core/src/cameras/camera.cuh:150-223): emission, background on a miss,
// material scatter with the light/cosine mixture pdf for Lambertian surfaces. Updates the path in
// place; returns true while the path continues (o, d then hold the next segment).
// SHAPES = false leaves out the analytic shapes' shading (the caller knows no scene of the pass has any).
template <int MODE, bool SHAPES>
__device__ __forceinline__ bool shade_hit(const crt_dev_scene* __restrict__ sc, const crt_dev_request& R, uint32_t pix,
                                          uint32_t sample, float4 hv, uint32_t iw, V3& o, V3& d, V3& att, V3& L,
                                          uint32_t& depth, uint32_t& bounce, Rng<MODE>& g, uint32_t& light_tests) {
    uint32_t prim = fbits(hv.y);
    bool passthrough = (iw & CRT_PASSTHROUGH) != 0;
    uint32_t inst = iw & CRT_INST_MASK;
    bool done = true;
    // a visible spot light passed through from behind emits colour(1,1,1,0) (diffuse_spot_light.cuh:
    // 41-44), added before whatever the re-trace found is shaded (camera.cuh:179-193)
    if (iw & CRT_PASSTHROUGH_WHITE) L = L + att * v3(1.0f, 1.0f, 1.0f);
    if (prim == CRT_MISS) {
        // miss: background (camera.cuh:161-174); a missed pass-through re-trace keeps the
        // light record, which does not scatter and adds colour(0,0,0,0)
        if (!passthrough) L = L + att * v3(sc->background[0], sc->background[1], sc->background[2]);
    } else {
        bounce += 1u;
        if (!g.minstd()) g.pcg = pcg_keyed(R.seed, ref_pixel_of(R, pix), sample, bounce);
        uint32_t type = CRT_HIT_TYPE(prim), idx = CRT_HIT_INDEX(prim);
        float t = hv.x;
        V3 p = o + t * d;
        V3 outward;
        uint32_t mat_id;
        float tu = hv.z, tvv = hv.w;  // texture coordinates (filled lazily below)
        bool need_uv_sphere = false;
        bool is_light_quad = false;
        bool shape = false, shape_front = false;
        V3 sphere_c; float sphere_r = 1.0f;
        uint32_t tri_flags = 0;
        float4 tq0, tq1, tq2;
        if (type == CRT_PRIM_TRI) {
            const float4* tp = (const float4*)sc->tris + 3 * idx;
            tq0 = ldg(tp); tq1 = ldg(tp + 1); tq2 = ldg(tp + 2);
            mat_id = fbits(tq1.w);
            tri_flags = fbits(tq2.w);
            // flat normal = unit(cross(v0 - v1, v0 - v2)) (core/src/primitives/triangle.cuh:118)
            V3 a = v3(tq0), b = v3(tq1), c = v3(tq2);
            V3 nrm = unit(cross(a - b, a - c));
            if (tri_flags & CRT_TF_SMOOTH) {
                const float4* tn = (const float4*)sc->tri_normals + 3 * idx;
                // u weights v1, v weights v2 (triangle.cuh:167-174), not renormalised
                nrm = hv.z * v3(ldg(tn + 1)) + hv.w * v3(ldg(tn + 2)) + (1.0f - hv.z - hv.w) * v3(ldg(tn));
            }
            if (inst != CRT_NO_INST) {
                // object -> world normal: transpose of the world -> object 3x3
                const float4* ip = (const float4*)sc->insts + 4 * inst;
                float4 r0 = ldg(ip), r1 = ldg(ip + 1), r2 = ldg(ip + 2);
                int ovr = (int)fbits(ldg(ip + 3).z);
                if (ovr >= 0) mat_id = (uint32_t)ovr;
                V3 nw = v3(r0.x * nrm.x + r1.x * nrm.y + r2.x * nrm.z, r0.y * nrm.x + r1.y * nrm.y + r2.y * nrm.z,
                           r0.z * nrm.x + r1.z * nrm.y + r2.z * nrm.z);
                nrm = (tri_flags & CRT_TF_SMOOTH) ? nw : unit(nw);
            }
            outward = nrm;
        } else if (type == CRT_PRIM_SPHERE) {
            float4 s = ldg((const float4*)sc->spheres + 2 * idx);
            mat_id = fbits(ldg((const float4*)sc->spheres + 2 * idx + 1).y);
            sphere_c = v3(s); sphere_r = s.w;
            outward = (p - sphere_c) / sphere_r;
            need_uv_sphere = true;
        } else {
            const float4* q = (const float4*)sc->quads + 5 * idx;
            float4 q1 = ldg(q + 1), q3 = ldg(q + 3), q4 = ldg(q + 4);
            mat_id = fbits(q4.w);
            const uint32_t qf = fbits(q1.w);
            outward = v3(q3);
            is_light_quad = (qf & CRT_QF_LIGHT) != 0;
            if (SHAPES && (qf & CRT_QF_SHAPE)) {
                shape = true;
                const float4 q0 = ldg(q), q2 = ldg(q + 2);
                V3 lo, ld;
                shape_local_ray(q0, q2, q3, q4, o, d, lo, ld);
                const V3 hp = lo + t * ld;   // the hit relative to the centre, in the shape's frame
                const float r = q1.x, h = q1.y;
                const float4* mp = (const float4*)sc->mats + fbits(q4.y);
                const float4 m0 = ldg(mp), m1 = ldg(mp + 1);   // (scale, offset), (repeat, -)
                V3 out_l;
                float su, sv;
                const float kPi = 3.14159265358979323846f;
                if (qf & CRT_QF_DISK) {
                    // normal +y (disk.cuh:72-73); get_disk_uv on hit - centre (uvmapping.cuh:306-343)
                    out_l = v3(0.0f, 1.0f, 0.0f);
                    su = fmodf((hp.x / (2.0f * r) + 0.5f) * m1.x, 1.0f);
                    sv = fmodf((hp.z / (2.0f * r) + 0.5f) * m1.y, 1.0f);
                } else if (qf & CRT_QF_CYLINDER) {
                    // (hit - centre) / radius with its y part, not renormalised (cylinder.cuh:93-95);
                    // get_cylinder_uv on it (uvmapping.cuh:285-304; M_PI is a float literal,
                    // core/src/misc/constants.cuh:35)
                    out_l = hp / r;
                    const float th = atan2f(out_l.x, out_l.z);
                    su = 1.0f - (th + kPi) / (2.0f * kPi);
                    sv = (out_l.y + h / 2.0f) / h;
                    su = fmodf(su * m1.x, 1.0f);
                    sv = fmodf(sv * m1.y, 1.0f);
                } else {
                    // cone.cuh:108-121: caps' normals at the end planes, else (x, |xz| r / h, z), unit;
                    // get_cone_uv on the vector before normalising (uvmapping.cuh:346-368, which
                    // applies the u repeat twice)
                    V3 ov;
                    if (hp.y <= 0.0f) ov = v3(0.0f, -1.0f, 0.0f);
                    else if (hp.y >= h) ov = v3(0.0f, 1.0f, 0.0f);
                    else ov = v3(hp.x, sqrtf(hp.x * hp.x + hp.z * hp.z) * (r / h), hp.z);
                    out_l = unit(ov);
                    const float th = atan2f(ov.x, ov.z);
                    su = (th + kPi) / (2.0f * kPi);
                    sv = sqrtf(ov.x * ov.x + ov.z * ov.z) / r;
                    su = fmodf(su * m1.x, 1.0f);
                    su = fmodf(su * m1.x, 1.0f);
                    sv = fmodf(sv * m1.y, 1.0f);
                }
                tu = m0.x * su + m0.z;
                tvv = m0.y * sv + m0.w;
                shape_front = dot(ld, out_l) < 0.0f;   // decided in the shape's frame (set_face_normal)
                // back to the world: A^T (rotation back, then divided by the scale), unit under a scale
                outward = v3(q0.x * out_l.x + q2.x * out_l.y + q3.x * out_l.z, q0.y * out_l.x + q2.y * out_l.y + q3.y * out_l.z,
                             q0.z * out_l.x + q2.z * out_l.y + q3.z * out_l.z);
                if (qf & CRT_QF_UNIT_NORMAL) outward = unit(outward);
            }
        }
        bool front = shape ? shape_front : dot(d, outward) < 0.0f;
        V3 nrm = front ? outward : -outward;
        const float4* mats = (const float4*)sc->mats;
        Mat m;
        m.m0 = ldg(mats + 2 * mat_id);
        m.m1 = ldg_i4(mats + 2 * mat_id + 1);
        if (m.m1.x == CRT_MAT_DIFFUSE_LIGHT) {
            // emitter: light does not scatter (material::scatter returns false)
            if (!passthrough) {
                bool one_sided = (m.m1.z & (int)CRT_MATF_ONE_SIDED) != 0 || is_light_quad;
                if (!one_sided || front) L = L + att * (v3(m.m0) * m.m0.w);
            }
        } else if (m.m1.x == CRT_MAT_SPOT_LIGHT) {
            // diffuse_spot_light::emitted (core/src/materials/diffuse_spot_light.cuh:38-61) on the front
            // face (back faces were passed through by the closest-hit stage): white outside the cone,
            // colour * intensity * cos^falloff inside; the cone axis (unit) and cos(cutoff) follow the
            // material record, then the falloff
            if (!passthrough) {
                const float4 sp0 = ldg(mats + 2 * mat_id + 2), sp1 = ldg(mats + 2 * mat_id + 3);
                const float ct = dot(unit(d), v3(sp0));
                const V3 e = ct < sp0.w ? v3(1.0f, 1.0f, 1.0f) : v3(m.m0) * (m.m0.w * powf(ct, sp1.x));
                L = L + att * e;
            }
        } else {
            // albedo from texture (lambertian/metal use the diffuse texture)
            V3 alb = v3(m.m0);
            if (m.m1.y >= 0) {
                if (need_uv_sphere) {
                    // get_sphere_uv on the outward normal (core/src/utilities/uvmapping.cuh:204-228)
                    const float2* mp = (const float2*)sc->sphere_map + 3 * idx;
                    float2 scl = ldg(mp), ofs = ldg(mp + 1), rep = ldg(mp + 2);
                    float theta = acosf(-outward.y);
                    float phi = atan2f(-outward.z, outward.x) + 3.14159265358979323846f;
                    float s_ = phi / (2.0f * 3.14159265358979323846f);
                    float t_ = theta / 3.14159265358979323846f;
                    s_ = fmodf(s_ * rep.x, 1.0f);
                    t_ = fmodf(t_ * rep.y, 1.0f);
                    tu = scl.x * s_ + ofs.x;
                    tvv = scl.y * t_ + ofs.y;
                } else if (type == CRT_PRIM_TRI) {
                    tu = 0.0f; tvv = 0.0f;
                    if (tri_flags & CRT_TF_UV) {
                        const float2* uvp = (const float2*)sc->tri_uvs + 3 * idx;
                        float2 a = ldg(uvp), b = ldg(uvp + 1), c = ldg(uvp + 2);
                        float w0 = 1.0f - hv.z - hv.w;
                        tu = w0 * a.x + hv.z * b.x + hv.w * c.x;
                        tvv = w0 * a.y + hv.z * b.y + hv.w * c.y;
                        tu = fmodf(tu, 1.0f); tvv = fmodf(tvv, 1.0f);
                        if (tu < 0.0f) tu += 1.0f;
                        if (tvv < 0.0f) tvv += 1.0f;
                    }
                }
                float4 tx = tex_value((const float4*)sc->texs, sc->texels, m.m1.y, tu, tvv, p);
                alb = v3(tx);
            }
            if (m.m1.x == CRT_MAT_METAL) {
                // metal::scatter (core/src/materials/metal.cuh:57-65): never absorbs
                V3 ud = unit(d);
                V3 refl = ud - (nrm * dot(nrm, ud)) * 2.0f;
                float x, y, z;
                do {
                    float a = g.uniform(), b = g.uniform(), c = g.uniform();
                    x = 2.0f * a - 1.0f; y = 2.0f * b - 1.0f; z = 2.0f * c - 1.0f;
                } while (x * x + y * y + z * z >= 1.0f);
                d = refl + m.m0.w * v3(x, y, z);
                att = att * alb;
                done = false;
            } else if (m.m1.x == CRT_MAT_DIELECTRIC) {
                // dielectric::scatter (core/src/materials/dielectric.cuh:32-51)
                float ir = m.m0.w;
                float ratio = front ? (1.0f / ir) : ir;
                V3 ud = unit(d);
                float dn = dot(-ud, nrm);
                float cos_t = dn < 1.0f ? dn : 1.0f;
                float sin_t = sqrtf(1.0f - cos_t * cos_t);
                bool reflect_ray = ratio * sin_t > 1.0f;
                if (!reflect_ray) {
                    float r0 = (1.0f - ratio) / (1.0f + ratio);
                    r0 = r0 * r0;
                    float x = 1.0f - cos_t;
                    float x5 = powf(x, 5.0f);
                    reflect_ray = (r0 + (1.0f - r0) * x5) > g.uniform();
                }
                if (reflect_ray) {
                    d = ud - (nrm * dot(nrm, ud)) * 2.0f;
                } else {
                    float dd = dot(nrm, ud);
                    float kk = 1.0f - ratio * ratio * (1.0f - dd * dd);
                    d = kk >= 0.0f ? (ratio * ud - (ratio * dd + sqrtf(kk)) * nrm) : v3(0.0f, 0.0f, 0.0f);
                }
                done = false;
            } else if ((m.m1.x == CRT_MAT_LAMBERTIAN || m.m1.x == CRT_MAT_ISOTROPIC || m.m1.x == CRT_MAT_OREN_NAYAR) &&
                       (m.m1.x != CRT_MAT_OREN_NAYAR || sc->num_lights > 0)) {
                // mixture of light sampling and cosine sampling (camera.cuh:203-216); isotropic
                // (isotropic.cuh) mixes the lights with sphere_pdf instead: uniform directions
                // (get_unit_vector, gpu_randomizer.cuh:53-60) and pdf and scattering pdf 1 / (4 pi)
                const bool iso = m.m1.x == CRT_MAT_ISOTROPIC;
                // oren_nayar::scatter (oren_nayar.cuh): a random_on_hemisphere direction it never uses
                // (two draws), sphere_pdf in the mixture, the cosine scattering pdf, and the
                // attenuation of light 0 (world light order) below
                const bool on = m.m1.x == CRT_MAT_OREN_NAYAR;
                if (on) { g.skip(); g.skip(); }
                const bool sph = iso || on;
                // onb::build_from_w (core/src/misc/onb.cuh:122-128)
                V3 w = unit(nrm);
                V3 aa = fabsf(w.x) > 0.9f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
                V3 bv = unit(cross(w, aa));
                V3 bu = cross(w, bv);
                const uint32_t nl = sc->num_lights;
                const float4* qd = (const float4*)sc->quads;
                const float4* sd = (const float4*)sc->spheres;
                V3 dir;
                if (g.uniform() < 0.5f) {
                    // hittable_list::random -> light 0 (get_int quirk) -> directional_light::random or,
                    // for a sphere light, omni_light::random (omni_light.cuh:90-98): random_to_sphere
                    // (gpu_randomizer.cuh:82-93) in the frame built on the direction to the centre
                    if (nl == 0) {
                        dir = v3(0.0f, 0.0f, 0.0f);
                    } else {
                        if (nl > 1) g.skip();
                        const uint32_t l0 = ldg(sc->lights);
                        if (l0 & CRT_LIGHT_SPHERE) {
                            const float4 s = ldg(sd + 2 * (l0 & ~CRT_LIGHT_SPHERE));
                            const V3 dc = v3(s) - p;
                            const float d2 = len2(dc);
                            const V3 lw = unit(dc);
                            const V3 la = fabsf(lw.x) > 0.9f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
                            const V3 lv = unit(cross(lw, la));
                            const V3 lu = cross(lw, lv);
                            const float r1 = g.uniform();
                            const float r2 = g.uniform();
                            const float z = 1.0f + r2 * (sqrtf(1.0f - s.w * s.w / d2) - 1.0f);
                            const float phi = 6.28318548202514648438f * r1;   // 2 * M_PI (float M_PI)
                            const float x = cosf(phi) * sqrtf(1.0f - z * z);
                            const float y = sinf(phi) * sqrtf(1.0f - z * z);
                            dir = x * lu + y * lv + z * lw;
                        } else {
                            const float4* q = qd + 5 * l0;
                            float a = g.uniform(), b = g.uniform();
                            V3 pp = v3(ldg(q)) + (a - 0.5f) * v3(ldg(q + 1)) + (b - 0.5f) * v3(ldg(q + 2));
                            dir = pp - p;
                        }
                    }
                } else if (sph) {
                    const float a = g.uniform() * 6.28318530718f;   // get_real(0, M_DOUBLE_PI)
                    const float z = g.uniform() * 2.0f - 1.0f;      // get_real(-1, 1)
                    const float r = sqrtf(1.0f - z * z);
                    dir = v3(r * cosf(a), r * sinf(a), z);
                } else {
                    // random_cosine_direction (core/src/misc/gpu_randomizer.cuh:122-133)
                    float r1 = g.uniform(), r2 = g.uniform();
                    float phi = 6.28318530718f * r1;
                    float z = sqrtf(1.0f - r2);
                    float rs = sqrtf(r2);
                    float x = cosf(phi) * rs, y = sinf(phi) * rs;
                    dir = x * bu + y * bv + z * w;
                }
                // light pdf (hittable_list::pdf_value -> directional_light::pdf_value, or for a
                // sphere light omni_light::pdf_value, omni_light.cuh:80-88: one over the cone's
                // solid angle 2 pi (1 - cos theta_max) when the direction meets the sphere)
                float weight = 1.0f / (float)nl;
                float sum = 0.0f;
                for (uint32_t li = 0; li < nl; ++li) {
                    const uint32_t lr = ldg(sc->lights + li);
                    float tl;
                    light_tests++;
                    float pv = 0.0f;
                    if (lr & CRT_LIGHT_SPHERE) {
                        const float4 s = ldg(sd + 2 * (lr & ~CRT_LIGHT_SPHERE));
                        if (sphere_hit(p, dir, s, 0.00001f, __builtin_inff(), tl)) {
                            const float cmax = sqrtf(1.0f - s.w * s.w / len2(v3(s) - p));
                            pv = 1.0f / (6.28318548202514648438f * (1.0f - cmax));
                        }
                    } else {
                        const float4* q = qd + 5 * lr;
                        float qa, qb;
                        uint32_t qkey;
                        if (quad_hit<false>(p, dir, q, 0.00001f, __builtin_inff(), false, tl, qa, qb, qkey)) {
                            float dist2 = tl * tl * len2(dir);
                            float cosine = fabsf(dot(dir, v3(ldg(q + 3))) / vlen(dir));
                            pv = dist2 / (cosine * ldg(q + 2).w);
                        }
                    }
                    sum += weight * pv;
                }
                V3 ud = unit(dir);
                float cv = dot(ud, w) / 3.14159265358979323846f;
                const float inv4pi = 1.0f / (4.0f * 3.14159265358979323846f);
                float pc = sph ? inv4pi : (0.0f > cv ? 0.0f : cv);
                float pdf = 0.5f * sum + (1.0f - 0.5f) * pc;
                float ct = dot(nrm, ud);
                float spdf = iso ? inv4pi : (ct < 0.0f ? 0.0f : ct / 3.14159265358979323846f);
                if (on) {
                    // light 0's colour * intensity * albedo * albedo_temp * (1 / pi) * (A + B alpha beta),
                    // the outgoing cosine 0 (skip_pdf_ray unset): sin theta_o = 1
                    const uint32_t l0 = ldg(sc->lights);
                    const uint32_t lmat = (l0 & CRT_LIGHT_SPHERE) ? fbits(ldg(sd + 2 * (l0 & ~CRT_LIGHT_SPHERE) + 1).y)
                                                                  : fbits(ldg(qd + 5 * l0 + 4).w);
                    const float4 lm = ldg(mats + 2 * lmat);
                    const float cosi = dot(-d, nrm);
                    const float mc = fmaxf(0.0f, cosi);
                    const float si = sqrtf(1.0f - mc * mc);
                    const float alpha = fmaxf(si, 1.0f) / fminf(si, 1.0f), beta = fminf(si, 1.0f) / fmaxf(si, 1.0f);
                    const float dr = m.m0.w + __int_as_float(m.m1.w) * alpha * beta;
                    alb = v3(lm) * lm.w * alb * __int_as_float(m.m1.z) * (1.0f / 3.14159265358979323846f) * dr;
                }
                att = att * ((1.0f / pdf) * (alb * spdf));
                d = dir;
                done = false;
            }
        }
        if (!done) {
            o = p;
            depth -= 1u;
            if (depth == 0u) done = true;
        }
    }
    return !done;
}

// Shade stage. FLAT (no instanced scene and no scene with analytic shapes among the paths) holds the
// kernel to 6 resident workgroups per CU (80 VGPRs, a few spills outside the hot paths), measured
// 2 % faster per frame on single-level scenes, and leaves the shapes' shading out; with instances
// (transforms, textures) the unbounded kernel is faster.
// ONE: the pass holds a single request, so the request record and its scene's pointers are
// wave-uniform (scalar loads once per wave instead of two dependent vector loads per path).
template <int MODE, int BLOCK, bool FLAT, bool ONE>
__global__ __launch_bounds__(BLOCK, FLAT ? 6 : 1) void k_shade(const crt_dev_scene* __restrict__ scenes,
                                                 const crt_dev_request* __restrict__ reqs,
                                                 const uint32_t* __restrict__ count_in, crt_paths Pin, crt_paths Pout,
                                                 uint32_t* __restrict__ count_out, const float4* __restrict__ hit,
                                                 const uint32_t* __restrict__ hit_inst, float4* __restrict__ out_rad,
                                                 uint32_t* __restrict__ live_min, int sort_octant,
                                                 crt_dev_counters* __restrict__ counters) {
    __shared__ uint32_t s_oct[8 * (BLOCK / 64)];
    __shared__ uint32_t s_base;
    __shared__ uint32_t s_min[BLOCK / 64];
    const uint32_t n = *count_in;
    uint32_t min_pid = 0xffffffffu;   // smallest path id among this thread's surviving paths
    const uint32_t stride = gridDim.x * BLOCK;
    uint32_t light_tests = 0;
    // grid-stride loop with a block-uniform trip count so every thread reaches the block scan
    const uint32_t iters = (n + stride - 1) / stride;
    for (uint32_t it = 0; it < iters; ++it) {
        uint32_t i = it * stride + blockIdx.x * BLOCK + threadIdx.x;
        bool keep = false;
        float4 no, nd, nt, nr;
        uint32_t oct = 0;   // direction octant of the next segment
        uint2 idv = make_uint2(0u, 0u);
        if (i < n) {
            float4 hv = hit[i];
            uint32_t iw = hit_inst[i];
            float4 ro = Pin.ray_o[i], rdv = Pin.ray_d[i];
            float4 tv = Pin.thr[i], rv = Pin.rad[i];
            idv = Pin.ids[i];
            uint32_t state = fbits(rdv.w);
            const uint32_t req = ONE ? 0u : (idv.y >> 16);
            const crt_dev_request& R = reqs[req];
            const crt_dev_scene* sc = scenes + R.scene;
            V3 o = v3(ro), d = v3(rdv);
            V3 att = v3(tv), L = v3(rv);
            uint32_t pix = idv.x, sample = idv.y & 0xffffu;
            uint32_t depth = ST_DEPTH(state), bounce = ST_BOUNCE(state);
            // closest-hit queries of this segment beyond the first: the pass-through re-trace
            const uint32_t retraces = ST_RETRACES(state) + ((iw & CRT_PASSTHROUGH) ? 1u : 0u);
            Rng<MODE> g;
            g.ms = fbits(tv.w);
            g.dyn = R.rng_mode;
            keep = shade_hit<MODE, !FLAT>(sc, R, pix, sample, hv, iw, o, d, att, L, depth, bounce, g, light_tests);
            if (keep) {
                min_pid = min(min_pid, R.path_base + (pix - R.pix_begin) * R.sample_count + (sample - R.sample_begin));
                uint32_t ns = (depth & 0xffu) | ((bounce & 0xffu) << 8) | (retraces << 24);
                no = make_float4(o.x, o.y, o.z, ro.w);
                nd = make_float4(d.x, d.y, d.z, bitsf(ns));
                nt = make_float4(att.x, att.y, att.z, bitsf(g.ms));
                nr = make_float4(L.x, L.y, L.z, 0.0f);
                oct = (d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u);
            } else {
                // finished: the radiance goes to the path's id, with the path's closest-hit queries
                // (one per hit, the final miss and every pass-through re-trace) in .w for the
                // per-request ray count
                uint32_t pid = R.path_base + (pix - R.pix_begin) * R.sample_count + (sample - R.sample_begin);
                const uint32_t rays = bounce + (fbits(hv.y) == CRT_MISS ? 1u : 0u) + retraces;
                out_rad[pid] = make_float4(L.x, L.y, L.z, (float)rays);
                if (g.minstd()) stg(R.rng_state + pix, g.ms);
            }
        }
        // survivors go to the front of the other state buffer (grouped by direction octant when
        // sort_octant is set; a launch-uniform choice)
        uint32_t pos = sort_octant ? block_append_sorted<BLOCK>(keep, oct, count_out, s_oct, &s_base)
                                   : block_append<BLOCK>(keep, count_out, s_oct, &s_base);
        if (keep) {
            Pout.ray_o[pos] = no;
            Pout.ray_d[pos] = nd;
            Pout.thr[pos] = nt;
            Pout.rad[pos] = nr;
            Pout.ids[pos] = idv;
        }
    }
    if (live_min) {
        // per block: the smallest live path id after this stage (0xffffffff: none). Every path id
        // below both this minimum over all blocks and the generated count has finished, which is
        // how the serving loop knows a request-table entry is complete (no per-path counters).
        uint32_t m = min_pid;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m = min(m, (uint32_t)__shfl_xor((int)m, off));
        if ((threadIdx.x & 63u) == 0u) s_min[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t b = s_min[0];
            for (int w = 1; w < BLOCK / 64; ++w) b = min(b, s_min[w]);
            live_min[blockIdx.x] = b;
        }
    }
    wave_add(&counters->light_tests, light_tests);
}

// One-render-per-launch path tracer, the reference's render kernel design (core/src/gpu.cu:756-801):
// each thread owns pixels, runs their samples in order and follows every path to its end, tracing
// and shading in one kernel. Same estimator, streams and per-pixel float sums as the wavefront
// pipeline, so the images agree; it serves as the plain baseline the batched passes are measured
// against. Grid-stride over the request's pixel range.
template <int MODE, int BLOCK, int RS, int LS, int GS, int WIDTH, int TL>
__global__ __launch_bounds__(BLOCK) void k_plain(const crt_dev_scene* __restrict__ scenes,
                                                 const crt_dev_request* __restrict__ req, uint32_t* __restrict__ gstack,
                                                 crt_dev_counters* __restrict__ counters) {
    __shared__ uint32_t s_stack[LS * BLOCK];
    const crt_dev_request& R = *req;
    const crt_dev_scene* sc = scenes + R.scene;
    const SceneView S = scene_view(sc);
    const uint32_t gtid = blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t stride = gridDim.x * BLOCK;
    uint32_t* lds_col = s_stack + threadIdx.x;
    const GSlot gst{gstack, gtid * GS};
    TravCount tc = {0, 0, 0, 0, 0, 0};
    uint32_t nrays = 0, nretrace = 0, novf = 0, light_tests = 0, nexact = 0;
    float4* accum = (float4*)R.accum;
    for (uint32_t q = gtid; q < R.pix_count; q += stride) {
        uint32_t pix = R.pix_begin + q;
        float4 acc = ldg(accum + pix);
        Rng<MODE> g;
        g.dyn = R.rng_mode;
        g.ms = 0u;
        if (g.minstd()) g.ms = ldg(R.rng_state + pix);
        const uint32_t refp = ref_pixel_of(R, pix);
        for (uint32_t k = 0; k < R.sample_count; ++k) {
            uint32_t sample = R.sample_begin + k;
            if (!g.minstd()) g.pcg = pcg_keyed(R.seed, refp, sample, 0u);
            V3 o, d;
            float tm;
            camera_ray<MODE>(R, pix, sample, g, o, d, tm);
            V3 att = v3(1.0f, 1.0f, 1.0f), L = v3(0.0f, 0.0f, 0.0f);
            uint32_t depth = R.max_depth & 0xffu, bounce = 0u;
            bool camera = true;
            for (;;) {
                HitRec h;
                const uint32_t pf = closest_hit_checked<BLOCK, RS, LS, GS, false, WIDTH, TL>(
                    S, o, d, camera, lds_col, gst, h, tc, novf, nrays, nretrace, nexact);
                camera = false;
                float4 hv = make_float4(h.t, bitsf(h.prim), h.u, h.v);
                if (!shade_hit<MODE, true>(sc, R, pix, sample, hv, h.inst | pf, o, d, att, L, depth, bounce, g, light_tests))
                    break;
            }
            acc.x += L.x; acc.y += L.y; acc.z += L.z;
        }
        acc.w += (float)R.sample_count;
        stg(accum + pix, acc);
        if (g.minstd()) stg(R.rng_state + pix, g.ms);
    }
    wave_add(&counters->rays, nrays);
    wave_add(&counters->retraces, nretrace);
    wave_add(&counters->light_tests, light_tests);
    wave_add(&counters->stack_overflow, novf);
    wave_add(&counters->exact, nexact);
}

// Per-pixel sums in sample order: accum += rad[sample 0] + ... (same float additions as the
// reference's pixel_color += ray_color(...) loop, core/src/gpu.cu:776-790). With `rays` set (one
// request), the paths' closest-hit queries (rad.w) are added to *rays: one atomic per wave.
// A pixel's samples are consecutive in rad, so a block stages 256 pixels x 8 samples at a time
// through LDS: the global loads are whole 128-byte lines (8 lanes per pixel) instead of one
// float4 per lane at a 16*spp-byte stride, and each lane still adds its pixel's samples in order.
constexpr uint32_t kResolveChunk = 8;
__global__ __launch_bounds__(256) void k_resolve(const crt_dev_request* __restrict__ reqs, int nreq,
                                                 const float4* __restrict__ rad, unsigned long long* __restrict__ rays) {
    __shared__ float4 stage[256 * (kResolveChunk + 1)];
    int r = blockIdx.y;
    if (r >= nreq) return;
    const crt_dev_request& R = reqs[r];
    const uint32_t S = R.sample_count;
    uint32_t nr = 0;
    // block-uniform trip counts: every lane reaches each barrier and the wave sum
    const uint32_t stride = gridDim.x * 256u;
    const uint32_t iters = (R.pix_count + stride - 1) / stride;
    for (uint32_t it = 0; it < iters; ++it) {
        const uint32_t q0 = it * stride + blockIdx.x * 256u;
        const uint32_t np = R.pix_count > q0 ? min(256u, R.pix_count - q0) : 0u;
        const uint32_t t = threadIdx.x;
        float4* a = (float4*)R.accum + R.pix_begin + q0 + t;
        float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (t < np) acc = ldg(a);
        const float4* src = rad + R.path_base + q0 * S;
        for (uint32_t k0 = 0; k0 < S; k0 += kResolveChunk) {
            const uint32_t n = min(kResolveChunk, S - k0);
            const uint32_t total = np * n;
            for (uint32_t e = t; e < total; e += 256u) {
                const uint32_t pl = (n == kResolveChunk) ? (e >> 3) : e / n;
                const uint32_t j = e - pl * n;
                stage[pl * (kResolveChunk + 1) + j] = src[pl * S + k0 + j];
            }
            __syncthreads();
            if (t < np) {
                const float4* row = stage + t * (kResolveChunk + 1);
                for (uint32_t j = 0; j < n; ++j) {
                    const float4 v = row[j];
                    acc.x += v.x; acc.y += v.y; acc.z += v.z;
                    nr += (uint32_t)v.w;
                }
            }
            __syncthreads();
        }
        if (t < np) {
            acc.w += (float)S;
            stg(a, acc);
        }
    }
    if (rays) wave_add(rays, nr);
}

// prepare_pixel_color + quantize (core/src/misc/color.cuh:289-309, core/src/gpu.cu:792-799).
__global__ __launch_bounds__(256) void k_tonemap(const float4* __restrict__ accum, uint32_t npix, uint32_t spp,
                                                 uint8_t* __restrict__ rgb) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    float4 a = accum[p];
    float c[3] = {a.x, a.y, a.z};
    for (int k = 0; k < 3; ++k) {
        float x = c[k];
        x = (x != x) ? 0.0f : x;
        if (spp > 0u) { float scale = 1.0f / (float)spp; x *= scale; }
        x = sqrtf(x);
        float cl = x < 0.0f ? 0.0f : (x > 0.999f ? 0.999f : x);
        rgb[3 * p + k] = (uint8_t)(int)(255.99f * cl);
    }
}

// Start of a request: zero its accumulation buffer and seed the per-pixel minstd engines with
// seed + reference pixel index (core/src/gpu.cu:770), Thrust's seed() rule (s mod m, 0 -> 1).
__global__ __launch_bounds__(256) void k_init_request(float4* __restrict__ accum, uint32_t* __restrict__ rng_state,
                                                      uint32_t W, uint32_t H, uint64_t seed) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= W * H) return;
    accum[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (rng_state) {
        uint32_t jj = p / W, i = p - jj * W;
        uint32_t ref_pixel = (H - 1u - jj) * W + i;
        uint64_t s = (seed + ref_pixel) & 0xffffffffull;
        uint32_t x = (uint32_t)(s % 2147483647ull);
        rng_state[p] = x == 0u ? 1u : x;
    }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// VALU issue-rate calibration: each lane runs `iters` rounds of 16 independent dependency chains
// of one instruction (KIND 0: v_fma_f32, 1: v_add_u32, 2: v_mul_f32), 16 wave instructions per
// round and wave, and writes its chains' sum so that none is removed. b and c come from the
// arguments, so nothing folds. KIND 3-7 run the v_fma_f32 rounds on part of each wave: lanes 0-15
// (one 16-lane group), every fourth lane (16 lanes over all four groups), lanes 0-31, lanes 0-62
// and lanes 0-47, which tells how a wave instruction's issue time depends on its active lanes.
// KIND 8 issues the 16 v_fma_f32 of a round with EXEC cleared (saved and restored inside one asm
// block, which touches no memory): what an instruction of a branch that no lane of the wave takes
// costs when the compiler does not jump around it. KIND 9-12 issue other instructions the traversal
// is made of, 16 per round: v_cndmask_b32 (selects of the child sort), v_max_f32 (slab tests),
// v_cmp_lt_f32 (compares into a lane mask) and v_mov_b32; KIND 13 the select in its three-operand
// form with the mask in an SGPR pair instead of VCC. KIND 14 and 15 issue the KIND 8 block with EXEC
// set to lanes 0-15 and to lanes 0-31 inside the asm block (every lane of the wave alive).
template <int KIND>
__global__ __launch_bounds__(256) void k_valu_probe(uint32_t iters, float b, float c, float* __restrict__ out) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    if ((KIND == 3 && lane >= 16u) || (KIND == 4 && (lane & 3u) != 0u) || (KIND == 5 && lane >= 32u) ||
        (KIND == 6 && lane == 63u) || (KIND == 7 && lane >= 48u)) {
        out[tid] = 0.0f;
        return;
    }
    if (KIND >= 9) {
        float a0 = (float)tid * 1e-7f, a1 = a0 + 1.0f, a2 = a0 + 2.0f, a3 = a0 + 3.0f;
        for (uint32_t i = 0; i < iters; ++i) {
            uint64_t m0, m1;
            if (KIND == 9)
                asm volatile(
                    "v_cmp_lt_f32 vcc, %[a0], %[a1]\n\t"
                    "v_cndmask_b32 %[a0], %[a0], %[a1], vcc\n\tv_cndmask_b32 %[a1], %[a1], %[a2], vcc\n\t"
                    "v_cndmask_b32 %[a2], %[a2], %[a3], vcc\n\tv_cndmask_b32 %[a3], %[a3], %[a0], vcc\n\t"
                    "v_cndmask_b32 %[a0], %[a0], %[a1], vcc\n\tv_cndmask_b32 %[a1], %[a1], %[a2], vcc\n\t"
                    "v_cndmask_b32 %[a2], %[a2], %[a3], vcc\n\tv_cndmask_b32 %[a3], %[a3], %[a0], vcc\n\t"
                    "v_cndmask_b32 %[a0], %[a0], %[a1], vcc\n\tv_cndmask_b32 %[a1], %[a1], %[a2], vcc\n\t"
                    "v_cndmask_b32 %[a2], %[a2], %[a3], vcc\n\tv_cndmask_b32 %[a3], %[a3], %[a0], vcc\n\t"
                    "v_cndmask_b32 %[a0], %[a0], %[a1], vcc\n\tv_cndmask_b32 %[a1], %[a1], %[a2], vcc\n\t"
                    "v_cndmask_b32 %[a2], %[a2], %[a3], vcc"
                    : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2), [a3] "+v"(a3) : : "vcc");
            else if (KIND == 10)
                asm volatile(
                    "v_max_f32 %[a0], %[a0], %[a1]\n\tv_max_f32 %[a1], %[a1], %[a2]\n\t"
                    "v_max_f32 %[a2], %[a2], %[a3]\n\tv_max_f32 %[a3], %[a3], %[a0]\n\t"
                    "v_max_f32 %[a0], %[a0], %[a1]\n\tv_max_f32 %[a1], %[a1], %[a2]\n\t"
                    "v_max_f32 %[a2], %[a2], %[a3]\n\tv_max_f32 %[a3], %[a3], %[a0]\n\t"
                    "v_max_f32 %[a0], %[a0], %[a1]\n\tv_max_f32 %[a1], %[a1], %[a2]\n\t"
                    "v_max_f32 %[a2], %[a2], %[a3]\n\tv_max_f32 %[a3], %[a3], %[a0]\n\t"
                    "v_max_f32 %[a0], %[a0], %[a1]\n\tv_max_f32 %[a1], %[a1], %[a2]\n\t"
                    "v_max_f32 %[a2], %[a2], %[a3]\n\tv_max_f32 %[a3], %[a3], %[a0]"
                    : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2), [a3] "+v"(a3));
            else if (KIND == 13)
                asm volatile(
                    "v_cmp_lt_f32 %[m0], %[a0], %[a1]\n\t"
                    "v_cndmask_b32_e64 %[a0], %[a0], %[a1], %[m0]\n\tv_cndmask_b32_e64 %[a1], %[a1], %[a2], %[m0]\n\t"
                    "v_cndmask_b32_e64 %[a2], %[a2], %[a3], %[m0]\n\tv_cndmask_b32_e64 %[a3], %[a3], %[a0], %[m0]\n\t"
                    "v_cndmask_b32_e64 %[a0], %[a0], %[a1], %[m0]\n\tv_cndmask_b32_e64 %[a1], %[a1], %[a2], %[m0]\n\t"
                    "v_cndmask_b32_e64 %[a2], %[a2], %[a3], %[m0]\n\tv_cndmask_b32_e64 %[a3], %[a3], %[a0], %[m0]\n\t"
                    "v_cndmask_b32_e64 %[a0], %[a0], %[a1], %[m0]\n\tv_cndmask_b32_e64 %[a1], %[a1], %[a2], %[m0]\n\t"
                    "v_cndmask_b32_e64 %[a2], %[a2], %[a3], %[m0]\n\tv_cndmask_b32_e64 %[a3], %[a3], %[a0], %[m0]\n\t"
                    "v_cndmask_b32_e64 %[a0], %[a0], %[a1], %[m0]\n\tv_cndmask_b32_e64 %[a1], %[a1], %[a2], %[m0]\n\t"
                    "v_cndmask_b32_e64 %[a2], %[a2], %[a3], %[m0]"
                    : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2), [a3] "+v"(a3), [m0] "=&s"(m0));
            else if (KIND == 11)
                asm volatile(
                    "v_cmp_lt_f32 %[m0], %[a0], %[a1]\n\tv_cmp_lt_f32 %[m1], %[a1], %[a2]\n\t"
                    "v_cmp_lt_f32 %[m0], %[a2], %[a3]\n\tv_cmp_lt_f32 %[m1], %[a3], %[a0]\n\t"
                    "v_cmp_lt_f32 %[m0], %[a0], %[a2]\n\tv_cmp_lt_f32 %[m1], %[a1], %[a3]\n\t"
                    "v_cmp_lt_f32 %[m0], %[a2], %[a0]\n\tv_cmp_lt_f32 %[m1], %[a3], %[a1]\n\t"
                    "v_cmp_lt_f32 %[m0], %[a0], %[a1]\n\tv_cmp_lt_f32 %[m1], %[a1], %[a2]\n\t"
                    "v_cmp_lt_f32 %[m0], %[a2], %[a3]\n\tv_cmp_lt_f32 %[m1], %[a3], %[a0]\n\t"
                    "v_cmp_lt_f32 %[m0], %[a0], %[a2]\n\tv_cmp_lt_f32 %[m1], %[a1], %[a3]\n\t"
                    "v_cmp_lt_f32 %[m0], %[a2], %[a0]\n\tv_cmp_lt_f32 %[m1], %[a3], %[a1]"
                    : [m0] "=&s"(m0), [m1] "=&s"(m1) : [a0] "v"(a0), [a1] "v"(a1), [a2] "v"(a2), [a3] "v"(a3));
            else
                asm volatile(
                    "v_mov_b32 %[a0], %[a1]\n\tv_mov_b32 %[a1], %[a2]\n\tv_mov_b32 %[a2], %[a3]\n\tv_mov_b32 %[a3], %[a0]\n\t"
                    "v_mov_b32 %[a0], %[a1]\n\tv_mov_b32 %[a1], %[a2]\n\tv_mov_b32 %[a2], %[a3]\n\tv_mov_b32 %[a3], %[a0]\n\t"
                    "v_mov_b32 %[a0], %[a1]\n\tv_mov_b32 %[a1], %[a2]\n\tv_mov_b32 %[a2], %[a3]\n\tv_mov_b32 %[a3], %[a0]\n\t"
                    "v_mov_b32 %[a0], %[a1]\n\tv_mov_b32 %[a1], %[a2]\n\tv_mov_b32 %[a2], %[a3]\n\tv_mov_b32 %[a3], %[a0]"
                    : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2), [a3] "+v"(a3));
            if (KIND == 11) a0 += (float)(uint32_t)(m0 ^ m1) * 0.0f;
            (void)m1;
        }
        out[tid] = a0 + a1 + a2 + a3;
        return;
    }
    if (KIND == 14 || KIND == 15) {
        float a0 = (float)tid * 1e-7f, a1 = a0 + 1.0f, a2 = a0 + 2.0f, a3 = a0 + 3.0f;
        const uint64_t em = KIND == 14 ? 0xffffull : 0xffffffffull;
        for (uint32_t i = 0; i < iters; ++i) {
            uint64_t sv;
            asm volatile(
                "s_mov_b64 %[sv], exec\n\t"
                "s_mov_b64 exec, %[em]\n\t"
                "v_fma_f32 %[a0], %[b], %[a0], %[c]\n\tv_fma_f32 %[a1], %[b], %[a1], %[c]\n\t"
                "v_fma_f32 %[a2], %[b], %[a2], %[c]\n\tv_fma_f32 %[a3], %[b], %[a3], %[c]\n\t"
                "v_fma_f32 %[a0], %[b], %[a0], %[c]\n\tv_fma_f32 %[a1], %[b], %[a1], %[c]\n\t"
                "v_fma_f32 %[a2], %[b], %[a2], %[c]\n\tv_fma_f32 %[a3], %[b], %[a3], %[c]\n\t"
                "v_fma_f32 %[a0], %[b], %[a0], %[c]\n\tv_fma_f32 %[a1], %[b], %[a1], %[c]\n\t"
                "v_fma_f32 %[a2], %[b], %[a2], %[c]\n\tv_fma_f32 %[a3], %[b], %[a3], %[c]\n\t"
                "v_fma_f32 %[a0], %[b], %[a0], %[c]\n\tv_fma_f32 %[a1], %[b], %[a1], %[c]\n\t"
                "v_fma_f32 %[a2], %[b], %[a2], %[c]\n\tv_fma_f32 %[a3], %[b], %[a3], %[c]\n\t"
                "s_mov_b64 exec, %[sv]"
                : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2), [a3] "+v"(a3), [sv] "=&s"(sv)
                : [b] "s"(b), [c] "v"(c), [em] "s"(em));
        }
        out[tid] = a0 + a1 + a2 + a3;
        return;
    }
    if (KIND == 8) {
        float a0 = (float)tid * 1e-7f, a1 = a0 + 1.0f, a2 = a0 + 2.0f, a3 = a0 + 3.0f;
        for (uint32_t i = 0; i < iters; ++i) {
            uint64_t sv;
            asm volatile(
                "s_mov_b64 %[sv], exec\n\t"
                "s_mov_b64 exec, 0\n\t"
                "v_fma_f32 %[a0], %[b], %[a0], %[c]\n\tv_fma_f32 %[a1], %[b], %[a1], %[c]\n\t"
                "v_fma_f32 %[a2], %[b], %[a2], %[c]\n\tv_fma_f32 %[a3], %[b], %[a3], %[c]\n\t"
                "v_fma_f32 %[a0], %[b], %[a0], %[c]\n\tv_fma_f32 %[a1], %[b], %[a1], %[c]\n\t"
                "v_fma_f32 %[a2], %[b], %[a2], %[c]\n\tv_fma_f32 %[a3], %[b], %[a3], %[c]\n\t"
                "v_fma_f32 %[a0], %[b], %[a0], %[c]\n\tv_fma_f32 %[a1], %[b], %[a1], %[c]\n\t"
                "v_fma_f32 %[a2], %[b], %[a2], %[c]\n\tv_fma_f32 %[a3], %[b], %[a3], %[c]\n\t"
                "v_fma_f32 %[a0], %[b], %[a0], %[c]\n\tv_fma_f32 %[a1], %[b], %[a1], %[c]\n\t"
                "v_fma_f32 %[a2], %[b], %[a2], %[c]\n\tv_fma_f32 %[a3], %[b], %[a3], %[c]\n\t"
                "s_mov_b64 exec, %[sv]"
                : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2), [a3] "+v"(a3), [sv] "=&s"(sv)
                : [b] "s"(b), [c] "v"(c));
        }
        out[tid] = a0 + a1 + a2 + a3;
        return;
    }
    if (KIND == 1) {
        uint32_t a[16];
        const uint32_t ib = __float_as_uint(b) | 1u;
#pragma unroll
        for (int k = 0; k < 16; ++k) a[k] = tid + (uint32_t)k;
        for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
            for (int k = 0; k < 16; ++k) a[k] = a[k] + ib;
        }
        uint32_t s = 0;
#pragma unroll
        for (int k = 0; k < 16; ++k) s ^= a[k];
        out[tid] = (float)s;
    } else {
        float a[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) a[k] = (float)(tid + (uint32_t)k) * 1e-7f;
        for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
            for (int k = 0; k < 16; ++k) a[k] = KIND == 2 ? a[k] * b : __builtin_fmaf(a[k], b, c);
        }
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < 16; ++k) s += a[k];
        out[tid] = s;
    }
}

// ======================================================================================== launchers
#define STK_RS 1
#define STK_LS 16
#define STK_GS 96

extern "C" int crt_k_pass_init(hipStream_t s, uint32_t* ctl, uint32_t nwords) {
    hipLaunchKernelGGL(k_pass_init, dim3(1), dim3(64), 0, s, ctl, nwords);
    return (int)hipGetLastError();
}

extern "C" int crt_k_regen(hipStream_t s, int rng_mode, uint32_t grid, const crt_dev_request* reqs, int nreq,
                           uint32_t pid0, uint32_t total, uint32_t nslots, uint32_t* ctl_in, uint32_t* ctl_out,
                           const uint32_t* open_pass, crt_paths P) {
#define CRT_RG1(M, O)                                                                                                \
    hipLaunchKernelGGL((k_regen<M, O>), dim3(grid), dim3(256), 0, s, reqs, nreq, pid0, total, nslots, ctl_in, ctl_out, \
                       open_pass, P)
#define CRT_RG(M) { if (nreq == 1 && !open_pass) CRT_RG1(M, true); else CRT_RG1(M, false); }
    if (rng_mode == CRT_RNG_MINSTD) CRT_RG(CRT_RNG_MINSTD)
    else if (rng_mode == CRT_RNG_PCG) CRT_RG(CRT_RNG_PCG)
    else CRT_RG(CRT_RNG_DYNAMIC)
#undef CRT_RG
#undef CRT_RG1
    return (int)hipGetLastError();
}

// Resident blocks of one kernel on this device (per-CU occupancy x CUs).
template <typename K> static uint32_t resident_blocks(K kernel, int block) {
    int nb = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, block, 0) != hipSuccess || nb <= 0) nb = 1;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    return (uint32_t)(nb * cus);
}

extern "C" int crt_k_extend(hipStream_t s, int count_mode, int width, int two_level, uint32_t grid,
                            const crt_dev_scene* scenes, const crt_dev_request* reqs, int32_t uniform_scene,
                            const uint32_t* queue, const uint32_t* count_ptr, uint32_t count_direct, const float4* ray_o,
                            const float4* ray_d, const uint2* ids, float4* hit, uint32_t* hit_inst, uint32_t* gstack,
                            uint32_t* zero_counter, int mode, uint32_t* work_counter, uint32_t param, int stack_lds,
                            const uint32_t* first_camera, crt_dev_counters* counters) {
    uint32_t us = uniform_scene >= 0 ? (uint32_t)uniform_scene : 0u;
    if (mode != CRT_EXTEND_WAVE_CHUNK) work_counter = nullptr;
    if (mode == CRT_EXTEND_WAVE_CHUNK && !work_counter) mode = CRT_EXTEND_GRID_STRIDE;
    // grid-stride / wave-chunk kernel with an LDS stack ring of LSV entries per lane
#define CRT_EXT_GS(LSV, C, U, W, T)                                                                                    \
    {                                                                                                                   \
        auto kfn = k_extend<256, STK_RS, LSV, STK_GS, C, U, W, T>;                                                     \
        static uint32_t rb = 0;                                                                                         \
        if (!rb) rb = resident_blocks(kfn, 256);                                                                        \
        uint32_t g = (mode == CRT_EXTEND_WAVE_CHUNK && grid > rb) ? rb : grid;                                          \
        hipLaunchKernelGGL(kfn, dim3(g), dim3(256), 0, s, scenes, reqs, us, queue, count_ptr, count_direct, ray_o,     \
                           ray_d, ids, hit, hit_inst, gstack, zero_counter, work_counter, param ? param : 1u,         \
                           counters);                                                                                   \
    }
    // lane-refill kernel over per-wave shares of the queue (same grid as the grid-stride kernel)
#define CRT_EXT_PT(C, U, W, T)                                                                                         \
    {                                                                                                                   \
        auto kfn = k_extend_pt<256, STK_RS, STK_LS, STK_GS, C, U, W, T>;                                               \
        hipLaunchKernelGGL(kfn, dim3(grid), dim3(256), 0, s, scenes, reqs, us, queue, count_ptr, count_direct, ray_o,  \
                           ray_d, ids, hit, hit_inst, gstack, zero_counter, param ? param : 1u, first_camera,       \
                           counters);                                                                                   \
    }
    // counting variants: one stack depth; timed variants: LDS ring of 8, 16 or 32 entries per lane
#define CRT_EXT_C(U, W, T) \
    if (mode == CRT_EXTEND_PERSISTENT) CRT_EXT_PT(true, U, W, T) else CRT_EXT_GS(STK_LS, true, U, W, T)
#define CRT_EXT_N(U, W, T)                                                     \
    if (mode == CRT_EXTEND_PERSISTENT) CRT_EXT_PT(false, U, W, T)              \
    else if (stack_lds == 8) CRT_EXT_GS(8, false, U, W, T)                     \
    else if (stack_lds == 32) CRT_EXT_GS(32, false, U, W, T)                   \
    else CRT_EXT_GS(STK_LS, false, U, W, T)
#define CRT_EXT_WT(X, U)                                                                     \
    if (width == 2) { if (two_level == 2) { X(U, 2, 2) } else if (two_level) { X(U, 2, 1) } else { X(U, 2, 0) } } \
    else { if (two_level == 2) { X(U, 4, 2) } else if (two_level) { X(U, 4, 1) } else { X(U, 4, 0) } }
    if (count_mode) {
        if (uniform_scene >= 0) { CRT_EXT_WT(CRT_EXT_C, true) } else { CRT_EXT_WT(CRT_EXT_C, false) }
    } else {
        if (uniform_scene >= 0) { CRT_EXT_WT(CRT_EXT_N, true) } else { CRT_EXT_WT(CRT_EXT_N, false) }
    }
#undef CRT_EXT_WT
#undef CRT_EXT_N
#undef CRT_EXT_C
#undef CRT_EXT_PT
#undef CRT_EXT_GS
    return (int)hipGetLastError();
}

extern "C" int crt_k_shade(hipStream_t s, int rng_mode, int full, uint32_t grid, const crt_dev_scene* scenes,
                           const crt_dev_request* reqs, const uint32_t* count_in, crt_paths Pin, crt_paths Pout,
                           uint32_t* count_out, const float4* hit, const uint32_t* hit_inst, float4* out_rad,
                           uint32_t* live_min, int sort_octant, int one_request, crt_dev_counters* counters) {
#define CRT_SH(M, F, O)                                                                                                   \
    hipLaunchKernelGGL((k_shade<M, 256, F, O>), dim3(grid), dim3(256), 0, s, scenes, reqs, count_in, Pin, Pout,         \
                       count_out, hit, hit_inst, out_rad, live_min, sort_octant, counters)
#define CRT_SH_O(M, F) { if (one_request) CRT_SH(M, F, true); else CRT_SH(M, F, false); }
#define CRT_SH_F(M) { if (full) CRT_SH_O(M, false) else CRT_SH_O(M, true) }
    if (rng_mode == CRT_RNG_MINSTD) CRT_SH_F(CRT_RNG_MINSTD)
    else if (rng_mode == CRT_RNG_PCG) CRT_SH_F(CRT_RNG_PCG)
    else CRT_SH_F(CRT_RNG_DYNAMIC)
#undef CRT_SH_F
#undef CRT_SH_O
#undef CRT_SH
    return (int)hipGetLastError();
}

extern "C" int crt_k_trace_rays(hipStream_t s, int width, int two_level, uint32_t grid, const crt_dev_scene* sc,
                                const float4* rays, uint32_t n, int camera, float4* hit, uint32_t* hit_inst,
                                uint32_t* gstack, crt_dev_counters* counters) {
#define CRT_TR(W, T)                                                                                              \
    hipLaunchKernelGGL((k_trace_rays<256, STK_RS, STK_LS, STK_GS, W, T>), dim3(grid), dim3(256), 0, s, sc, rays, n, \
                       camera, hit, hit_inst, gstack, counters)
    if (width == 2) { if (two_level == 2) CRT_TR(2, 2); else if (two_level) CRT_TR(2, 1); else CRT_TR(2, 0); }
    else { if (two_level == 2) CRT_TR(4, 2); else if (two_level) CRT_TR(4, 1); else CRT_TR(4, 0); }
#undef CRT_TR
    return (int)hipGetLastError();
}

extern "C" int crt_k_plain(hipStream_t s, int rng_mode, int width, int two_level, uint32_t grid,
                           const crt_dev_scene* scenes, const crt_dev_request* req, uint32_t* gstack,
                           crt_dev_counters* counters) {
#define CRT_PL(M, W, T)                                                                                          \
    hipLaunchKernelGGL((k_plain<M, 256, STK_RS, STK_LS, STK_GS, W, T>), dim3(grid), dim3(256), 0, s, scenes, req, \
                       gstack, counters)
#define CRT_PL_WT(M)                                                              \
    if (width == 2) { if (two_level == 2) CRT_PL(M, 2, 2); else if (two_level) CRT_PL(M, 2, 1); else CRT_PL(M, 2, 0); } \
    else { if (two_level == 2) CRT_PL(M, 4, 2); else if (two_level) CRT_PL(M, 4, 1); else CRT_PL(M, 4, 0); }
    if (rng_mode == CRT_RNG_MINSTD) { CRT_PL_WT(CRT_RNG_MINSTD) } else { CRT_PL_WT(CRT_RNG_PCG) }
#undef CRT_PL_WT
#undef CRT_PL
    return (int)hipGetLastError();
}

extern "C" int crt_k_resolve(hipStream_t s, const crt_dev_request* reqs, int nreq, uint32_t max_pix, const float4* rad,
                             unsigned long long* rays) {
    if (nreq == 0) return 0;
    uint32_t gx = (max_pix + 255) / 256;
    if (gx > 4096) gx = 4096;
    if (gx == 0) gx = 1;
    hipLaunchKernelGGL(k_resolve, dim3(gx, nreq), dim3(256), 0, s, reqs, nreq, rad, nreq == 1 ? rays : nullptr);
    return (int)hipGetLastError();
}

extern "C" int crt_k_tonemap(hipStream_t s, const float4* accum, uint32_t npix, uint32_t spp, uint8_t* rgb) {
    if (npix == 0) return 0;
    hipLaunchKernelGGL(k_tonemap, dim3((npix + 255) / 256), dim3(256), 0, s, accum, npix, spp, rgb);
    return (int)hipGetLastError();
}

extern "C" int crt_k_init_request(hipStream_t s, float4* accum, uint32_t* rng_state, uint32_t W, uint32_t H, uint64_t seed) {
    uint32_t n = W * H;
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_init_request, dim3((n + 255) / 256), dim3(256), 0, s, accum, rng_state, W, H, seed);
    return (int)hipGetLastError();
}

extern "C" int crt_k_fill_u32(hipStream_t s, uint32_t* p, uint32_t n, uint32_t v) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
    return (int)hipGetLastError();
}

extern "C" int crt_k_valu_probe(hipStream_t s, int kind, uint32_t blocks, uint32_t iters, float* out) {
    if (kind == 0) hipLaunchKernelGGL(k_valu_probe<0>, dim3(blocks), dim3(256), 0, s, iters, 0.999f, 1e-6f, out);
    else if (kind == 1) hipLaunchKernelGGL(k_valu_probe<1>, dim3(blocks), dim3(256), 0, s, iters, 0.999f, 1e-6f, out);
    else if (kind == 2) hipLaunchKernelGGL(k_valu_probe<2>, dim3(blocks), dim3(256), 0, s, iters, 0.999f, 1e-6f, out);
    else if (kind == 3) hipLaunchKernelGGL(k_valu_probe<3>, dim3(blocks), dim3(256), 0, s, iters, 0.999f, 1e-6f, out);
    else if (kind == 4) hipLaunchKernelGGL(k_valu_probe<4>, dim3(blocks), dim3(256), 0, s, iters, 0.999f, 1e-6f, out);
    else if (kind == 5) hipLaunchKernelGGL(k_valu_probe<5>, dim3(blocks), dim3(256), 0, s, iters, 0.999f, 1e-6f, out);
    else if (kind == 6) hipLaunchKernelGGL(k_valu_probe<6>, dim3(blocks), dim3(256), 0, s, iters, 0.999f, 1e-6f, out);
    else if (kind == 7) hipLaunchKernelGGL(k_valu_probe<7>, dim3(blocks), dim3(256), 0, s, iters, 0.999f, 1e-6f, out);
    else if (kind == 8) hipLaunchKernelGGL(k_valu_probe<8>, dim3(blocks), dim3(256), 0, s, iters, 0.999f, 1e-6f, out);
    else if (kind == 9) hipLaunchKernelGGL(k_valu_probe<9>, dim3(blocks), dim3(256), 0, s, iters, 0.999f, 1e-6f, out);
    else if (kind == 10) hipLaunchKernelGGL(k_valu_probe<10>, dim3(blocks), dim3(256), 0, s, iters, 0.999f, 1e-6f, out);
    else if (kind == 11) hipLaunchKernelGGL(k_valu_probe<11>, dim3(blocks), dim3(256), 0, s, iters, 0.999f, 1e-6f, out);
    else if (kind == 12) hipLaunchKernelGGL(k_valu_probe<12>, dim3(blocks), dim3(256), 0, s, iters, 0.999f, 1e-6f, out);
    else if (kind == 13) hipLaunchKernelGGL(k_valu_probe<13>, dim3(blocks), dim3(256), 0, s, iters, 0.999f, 1e-6f, out);
    else if (kind == 14) hipLaunchKernelGGL(k_valu_probe<14>, dim3(blocks), dim3(256), 0, s, iters, 0.999f, 1e-6f, out);
    else hipLaunchKernelGGL(k_valu_probe<15>, dim3(blocks), dim3(256), 0, s, iters, 0.999f, 1e-6f, out);
    return (int)hipGetLastError();
}

extern "C" uint32_t crt_k_stack_capacity(void) { return STK_RS + STK_LS + STK_GS; }
extern "C" uint32_t crt_k_gstack_entries(void) { return STK_GS; }
