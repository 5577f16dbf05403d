// oracle.cpp - CPU restatement of the reference render path. TEST INFRASTRUCTURE ONLY.
//
// This file is the checker, never the product: only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load it. It restates the reference's per-pixel estimator
// (core/src/gpu.cu:761-802 + core/src/cameras/camera.cuh:149-264) in two configurations:
//
//   refcpu   : Real = float, RNG = minstd (Thrust minstd_rand + uniform_real_distribution<float>,
//              core/src/gpu.cu:770-771, core/src/misc/gpu_randomizer.cuh:11-17), reference-style
//              BVH (random axis from minstd_rand(1984), sort on bbox min, count-median split,
//              core/src/misc/bvh_node.cuh:63-118), Moller-Trumbore triangles with |det|<1e-6 reject
//              (core/src/primitives/triangle.cuh:130-185), the reference's diagonal-only quad
//              boxes and stale triangle UV point. This is the "reference render" and the timed
//              CPU baseline.
//   oracle64 : Real = double, RNG = the keyed PCG32 streams of the product (seed, pixel, sample,
//              bounce), watertight triangle test, exact quad boxes. Same estimator otherwise.
//              This is the FP64 oracle for image RMSE and first-hit parity.
//
// Every function cites the reference file:line it follows. Pinning (tests/test_oracle.py):
// minstd against std::minstd_rand and the Thrust-documented 10000th output 399268537; PCG32
// against the vendored pcg_random.hpp (oracle/_ref, golden vectors in tests/golden/);
// intersections against closed forms and brute force; the estimator against the white-furnace
// closed form and Monte-Carlo normalisation of the light PDF; tonemap exhaustively.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <thread>
#include <vector>

#include "../include/cortex_rt.h"

namespace orc {

// ------------------------------------------------------------------ vector math (glm fvec3 semantics)
template <class R> struct V3 {
    R x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    R& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
// glm dot: tmp = a*b; tmp.x + tmp.y + tmp.z
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class R> inline R len2(V3<R> v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
template <class R> inline R length(V3<R> v) { return std::sqrt(len2(v)); }
// unit_vector: v / vector3(len, len, len) (core/src/misc/vector3.cuh:73-76)
template <class R> inline V3<R> unit(V3<R> v) { R l = length(v); return {v.x / l, v.y / l, v.z / l}; }
// glm::reflect: I - N * dot(N, I) * 2 (libs/glm/detail/func_geometric.inl:104-110)
template <class R> inline V3<R> reflect(V3<R> I, V3<R> N) { R d = dot(N, I); return I - (N * d) * R(2); }
// glm::refract (libs/glm/detail/func_geometric.inl:113-122)
template <class R> inline V3<R> refract(V3<R> I, V3<R> N, R eta) {
    R d = dot(N, I);
    R k = R(1) - eta * eta * (R(1) - d * d);
    if (!(k >= R(0))) return V3<R>();
    return eta * I - (eta * d + std::sqrt(k)) * N;
}
template <class R> inline V3<R> vmin(V3<R> a, V3<R> b) { return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)}; }
template <class R> inline V3<R> vmax(V3<R> a, V3<R> b) { return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }
template <class R> inline V3<R> from3(const float* p) { return {R(p[0]), R(p[1]), R(p[2])}; }

// ------------------------------------------------------------------ random streams
// Thrust minstd_rand: x <- 48271 x mod (2^31-1); seed(s): s mod m, 0 -> 1
// (/opt/rocm/include/thrust/random/detail/linear_congruential_engine.inl:41-50).
struct Minstd {
    uint32_t x;
    explicit Minstd(uint32_t s) { x = s % 2147483647u; if (x == 0) x = 1; }
    uint32_t next() { x = (uint32_t)((uint64_t)x * 48271u % 2147483647u); return x; }
    // uniform_real_distribution<float>(0,1): float(x - min) / (1 + float(max - min)), then *1 + 0
    // (/opt/rocm/include/thrust/random/detail/uniform_real_distribution.inl:67-80). The
    // denominator rounds to 2^31 in float.
    float uniform_f() {
        float r = (float)(next() - 1u);
        r /= (1.0f + (float)(2147483646u - 1u));
        return r * (1.0f - 0.0f) + 0.0f;
    }
    // uniform_int_distribution<int>(a,b) maps through uniform_real_distribution<double>(a, b+1)
    // (/opt/rocm/include/thrust/random/detail/uniform_int_distribution.inl:66-86).
    int uniform_int(int a, int b) {
        double r = (double)(next() - 1u);
        r /= (1.0 + (double)(2147483646u - 1u));
        return (int)(r * ((double)b + 1.0 - (double)a) + (double)a);
    }
};

// PCG32 (O'Neill, pcg-c-basic pcg32_random_r / pcg32_srandom_r; restated from the published
// algorithm, pinned against the vendored libs/pcg/pcg_random.hpp pcg32 engine).
struct Pcg32 {
    uint64_t state, inc;
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    void seed(uint64_t initstate, uint64_t initseq) {
        state = 0u; inc = (initseq << 1u) | 1u; next(); state += initstate; next();
    }
};
inline uint64_t splitmix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
// The product's keyed stream: one PCG32 stream per (seed, pixel, sample, bounce).
inline void keyed_seed(Pcg32& g, uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce) {
    uint64_t key = ((uint64_t)pixel << 32) | (uint64_t)sample;
    g.seed(splitmix64(seed ^ splitmix64(key)), (uint64_t)bounce);
}

// A draw source for one path: minstd (carried across samples of one pixel) or keyed PCG
// (re-keyed at every bounce). uniform() returns the same value in float and double oracles.
struct Rng {
    int mode = CRT_RNG_PCG;
    Minstd ms{1};
    Pcg32 pcg{};
    uint64_t seed = 0; uint32_t pixel = 0, sample = 0;
    void bounce(uint32_t b) { if (mode == CRT_RNG_PCG) keyed_seed(pcg, seed, pixel, sample, b); }
    float uniform() {
        if (mode == CRT_RNG_MINSTD) return ms.uniform_f();
        return (float)(pcg.next() >> 8) * (1.0f / 16777216.0f);
    }
    int uniform_int_min(int a, int b) {
        // get_int uses uniform_real_distribution<int>: integer division makes it always return a
        // (core/src/misc/gpu_randomizer.cuh:29-36); it still consumes one draw.
        (void)b;
        if (mode == CRT_RNG_MINSTD) ms.next(); else pcg.next();
        return a;
    }
};

// ------------------------------------------------------------------ scene
enum ObjType { OBJ_TRI = 0, OBJ_SPHERE = 1, OBJ_QUAD = 2, OBJ_LIGHT = 3, OBJ_SPHERE_LIGHT = 4, OBJ_SHAPE = 5 };

// The reference's translate / rotate / scale wrappers around one primitive
// (core/src/scene_factory.cuh:710-725: translate innermost, scale outermost). A world ray reaches the
// primitive divided by the scale (scale.cuh:50-56), rotated by Rx(-x) Ry(-y) Rz(-z) (glm::rotate
// post-multiplies, rotate.cuh:104-119) and offset by -T (translate.cuh:38-43). The normal goes back
// through the rotation's inverse Rz(z) Ry(y) Rx(x) and then the scale, where it is divided by the
// scale and made unit (scale.cuh:58-60). (The reference rotates the normal back by Rx(x) Ry(y) Rz(z),
// which is the same matrix for a rotation about one axis; this restatement takes the inverse for
// two or more axes, as the builder maps the geometry of the other primitives.)
template <class R> struct Xf {
    bool t = false, rot = false, sc = false;
    V3<R> T, S;
    R Ri[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};   // Rx(-x) Ry(-y) Rz(-z)
    void set(const crt_transform& x) {
        t = (x.has & CRT_XF_TRANSLATE) != 0; rot = (x.has & CRT_XF_ROTATE) != 0; sc = (x.has & CRT_XF_SCALE) != 0;
        T = V3<R>(R(x.translate[0]), R(x.translate[1]), R(x.translate[2]));
        S = V3<R>(R(x.scale[0]), R(x.scale[1]), R(x.scale[2]));
        if (!rot) return;
        for (int axis = 0; axis < 3; ++axis) {
            R a = -R(x.rotate_deg[axis]) * R(3.14159265358979323846) / R(180), c = std::cos(a), sn = std::sin(a);
            R M[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
            int i = (axis + 1) % 3, j = (axis + 2) % 3;
            M[3 * i + i] = c; M[3 * i + j] = -sn; M[3 * j + i] = sn; M[3 * j + j] = c;
            R P[9];
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q) P[3 * p + q] = Ri[3 * p] * M[q] + Ri[3 * p + 1] * M[3 + q] + Ri[3 * p + 2] * M[6 + q];
            for (int k = 0; k < 9; ++k) Ri[k] = P[k];
        }
    }
    V3<R> inv_rot(V3<R> v) const {
        return {Ri[0] * v.x + Ri[1] * v.y + Ri[2] * v.z, Ri[3] * v.x + Ri[4] * v.y + Ri[5] * v.z, Ri[6] * v.x + Ri[7] * v.y + Ri[8] * v.z};
    }
    V3<R> fwd_rot(V3<R> v) const {
        return {Ri[0] * v.x + Ri[3] * v.y + Ri[6] * v.z, Ri[1] * v.x + Ri[4] * v.y + Ri[7] * v.z, Ri[2] * v.x + Ri[5] * v.y + Ri[8] * v.z};
    }
    void to_object(V3<R>& o, V3<R>& d) const {
        if (sc) { o = V3<R>(o.x / S.x, o.y / S.y, o.z / S.z); d = V3<R>(d.x / S.x, d.y / S.y, d.z / S.z); }
        if (rot) { o = inv_rot(o); d = inv_rot(d); }
        if (t) o = o - T;
    }
    V3<R> normal_to_world(V3<R> n) const {
        if (rot) n = fwd_rot(n);
        if (sc) n = unit(V3<R>(n.x / S.x, n.y / S.y, n.z / S.z));
        return n;
    }
    V3<R> point_to_world(V3<R> p) const {
        if (t) p = p + T;
        if (rot) p = fwd_rot(p);
        if (sc) p = V3<R>(p.x * S.x, p.y * S.y, p.z * S.z);
        return p;
    }
};

template <class R> struct Box {
    V3<R> lo, hi;  // interval per axis
};

template <class R> struct Obj {
    int type;
    int32_t gid;       // global primitive id (lights, spheres, quads, triangles)
    int32_t inst;      // instance index or -1
    uint32_t mat;
    // triangle
    V3<R> v[3], n[3]; R uv[3][2]; bool smooth = false, has_uv = false;
    V3<R> e1, e2, middle_normal;
    // sphere
    V3<R> c; R r; crt_uvmap map;
    // quad / light
    V3<R> q, qu, qv, qn, qw; R qd, area; bool invisible = false;
    // analytic shape (CRT_SHAPE_*; centre and radius in c, r, mapping in map) under its wrappers
    uint32_t kind = 0; R h = 0; Xf<R> xf;
    // sphere light (centre and radius in c, r): emission, spot cone
    V3<R> le; R lint = 0; bool spot = false; V3<R> sdir; R scut = 0, sfall = 0;
    Box<R> box;
};

template <class R> struct Light {  // elights entry (core/src/gpu.cu:681-694)
    int obj;
};

template <class R> struct Mat { crt_material m; };

template <class R> struct Scene {
    std::vector<Obj<R>> objs;      // world list order: lights, spheres, quads, triangles
    std::vector<int> lights;       // indices into objs
    std::vector<crt_material> mats;
    std::vector<crt_texture> texs;
    std::vector<std::vector<uint8_t>> tex_pixels;
    V3<R> background;
    // camera (core/src/cameras/perspective_camera.cuh:41-110, orthographic_camera.cuh:57-142)
    V3<R> origin, llc, horizontal, vertical, cu, cv, cw;
    R lens_radius, time0, time1;
    bool ortho = false;
    // BVH (reference-style): nodes over objs
    struct Node { Box<R> box; int left, right; bool leaf_l, leaf_r; };
    std::vector<Node> nodes;
    int root = -1; bool root_is_leaf = false;
    bool ref_style = true;   // refcpu semantics (diagonal quad boxes, MT triangles, stale uv)
    int accel = 0;           // 0 = BVH, 1 = brute force
};

template <class R> inline Box<R> box_points(V3<R> a, V3<R> b) {
    return {vmin(a, b), vmax(a, b)};
}
// aabb::pad (core/src/misc/aabb.cuh:120-130)
template <class R> inline Box<R> pad(Box<R> b) {
    const R delta = R(0.0001f);
    for (int a = 0; a < 3; ++a) {
        if (!(b.hi[a] - b.lo[a] >= delta)) { R p = delta / R(2); b.lo[a] = b.lo[a] - p; b.hi[a] = b.hi[a] + p; }
    }
    return b;
}
template <class R> inline Box<R> merge(Box<R> a, Box<R> b) { return {vmin(a.lo, b.lo), vmax(a.hi, b.hi)}; }

template <class R> static V3<R> xform_point(const float* m, V3<R> p) {
    return {R(m[0]) * p.x + R(m[1]) * p.y + R(m[2]) * p.z + R(m[3]),
            R(m[4]) * p.x + R(m[5]) * p.y + R(m[6]) * p.z + R(m[7]),
            R(m[8]) * p.x + R(m[9]) * p.y + R(m[10]) * p.z + R(m[11])};
}
template <class R> static V3<R> xform_vec(const float* m, V3<R> p) {
    return {R(m[0]) * p.x + R(m[1]) * p.y + R(m[2]) * p.z,
            R(m[4]) * p.x + R(m[5]) * p.y + R(m[6]) * p.z,
            R(m[8]) * p.x + R(m[9]) * p.y + R(m[10]) * p.z};
}

template <class R> static void add_triangle(Scene<R>& s, V3<R> a, V3<R> b, V3<R> c, const float* na, const float* nb,
                                            const float* nc, const float* ta, const float* tb, const float* tc,
                                            uint32_t mat, int32_t gid, int32_t inst) {
    Obj<R> o{};
    o.type = OBJ_TRI; o.gid = gid; o.inst = inst; o.mat = mat;
    o.v[0] = a; o.v[1] = b; o.v[2] = c;
    if (na) {
        o.smooth = true;
        // vert_normals = unit_vector(vn) (core/src/primitives/triangle.cuh:97-99)
        o.n[0] = unit(from3<R>(na)); o.n[1] = unit(from3<R>(nb)); o.n[2] = unit(from3<R>(nc));
    }
    if (ta) {
        o.has_uv = true;
        o.uv[0][0] = ta[0]; o.uv[0][1] = ta[1]; o.uv[1][0] = tb[0]; o.uv[1][1] = tb[1]; o.uv[2][0] = tc[0]; o.uv[2][1] = tc[1];
    }
    o.e1 = b - a; o.e2 = c - a;
    // middle_normal = unit_vector(cross(v0 - v1, v0 - v2)) (core/src/primitives/triangle.cuh:118)
    o.middle_normal = unit(cross(a - b, a - c));
    V3<R> lo = vmin(vmin(a, b), c), hi = vmax(vmax(a, b), c);
    if (s.ref_style) {
        // bbox padded by 0.001 (core/src/primitives/triangle.cuh:121-127)
        V3<R> e(R(0.001f), R(0.001f), R(0.001f));
        o.box = box_points(lo - e, hi + e);
    } else {
        o.box = pad(Box<R>{lo, hi});  // exact box, widened only when flat
    }
    s.objs.push_back(o);
}

template <class R> static int build_scene(Scene<R>& s, const crt_scene_desc* d, bool ref_style, int accel) {
    s.ref_style = ref_style; s.accel = accel;
    s.mats.assign(d->materials, d->materials + d->num_materials);
    for (auto& m : s.mats) if (m.type == CRT_MAT_METAL) m.fuzz = m.fuzz < 1 ? m.fuzz : 1;  // metal ctor clamp
    s.texs.assign(d->textures, d->textures + d->num_textures);
    s.tex_pixels.resize(d->num_textures);
    for (uint32_t i = 0; i < d->num_textures; ++i) {
        const crt_texture& t = d->textures[i];
        if (t.type == CRT_TEX_IMAGE && t.pixels) {
            size_t n = (size_t)t.width * t.height * t.channels;
            s.tex_pixels[i].assign(t.pixels, t.pixels + n);
        }
    }
    s.background = from3<R>(d->background);
    int32_t gid = 0;
    // sphere lights: omni_light / spot_light ctors (core/src/lights/omni_light.cuh:22-32,
    // spot_light.cuh:40-50): box centre -/+ radius, the spot's cutoff stored as
    // cos(degrees_to_radians(cutoff)) in float, diffuse_spot_light normalises the direction per hit
    auto add_sphere_light = [&](const crt_sphere_light& L) {
        Obj<R> o{};
        o.type = OBJ_SPHERE_LIGHT; o.gid = gid++; o.inst = -1; o.mat = 0xffffffffu;
        o.c = from3<R>(L.center); o.r = R(L.radius);
        o.map = crt_uvmap{1.0f, 1.0f, 0.0f, 0.0f, 1.0f, 1.0f};
        o.invisible = L.invisible != 0;
        o.le = from3<R>(L.rgb); o.lint = R(L.intensity);
        o.spot = L.kind == CRT_LIGHT_SPOT;
        if (o.spot) {
            o.scut = R(std::cos(L.cutoff_deg * 3.14159265358979323846f / 180.0f));
            o.sdir = unit(from3<R>(L.direction));
            o.sfall = R(L.falloff);
        }
        V3<R> rv(o.r, o.r, o.r);
        o.box = box_points(o.c - rv, o.c + rv);
        s.objs.push_back(o);
        s.lights.push_back((int)s.objs.size() - 1);
    };
    // lights first (core/src/gpu.cu:504-523: omni, directional, spot), then spheres (:527-535),
    // quads (:548-556), meshes (:621-640)
    for (uint32_t i = 0; i < d->num_sphere_lights; ++i)
        if (d->sphere_lights[i].kind == CRT_LIGHT_OMNI) add_sphere_light(d->sphere_lights[i]);
    for (uint32_t i = 0; i < d->num_lights; ++i, ++gid) {
        const crt_area_light& L = d->lights[i];
        Obj<R> o{};
        o.type = OBJ_LIGHT; o.gid = gid; o.inst = -1; o.mat = 0xffffffffu;
        o.q = from3<R>(L.center); o.qu = from3<R>(L.u); o.qv = from3<R>(L.v);
        // directional_light ctor (core/src/lights/directional_light.cuh:45-58)
        V3<R> n = cross(o.qu, o.qv);
        o.qn = unit(n); o.qd = dot(o.qn, o.q); o.qw = n / dot(n, n); o.area = length(n);
        o.invisible = L.invisible != 0;
        o.c = from3<R>(L.rgb); o.r = R(L.intensity);  // emission colour / intensity
        if (ref_style) {
            // set_bounding_box: corners position -/+ 0.5(u+v), padded (directional_light.cuh:61-69)
            V3<R> h = R(0.5f) * (o.qu + o.qv);
            o.box = pad(box_points(o.q - h, o.q + h));
        } else {
            V3<R> c0 = o.q - R(0.5) * o.qu - R(0.5) * o.qv;
            Box<R> b = box_points(c0, c0 + o.qu);
            b = merge(b, box_points(c0 + o.qv, c0 + o.qu + o.qv));
            o.box = pad(b);
        }
        s.objs.push_back(o);
        s.lights.push_back((int)s.objs.size() - 1);
    }
    for (uint32_t i = 0; i < d->num_sphere_lights; ++i)
        if (d->sphere_lights[i].kind == CRT_LIGHT_SPOT) add_sphere_light(d->sphere_lights[i]);
    for (uint32_t i = 0; i < d->num_spheres; ++i, ++gid) {
        const crt_sphere& S = d->spheres[i];
        if (S.material >= d->num_materials) return CRT_ERR_INVALID_SCENE;
        Obj<R> o{};
        o.type = OBJ_SPHERE; o.gid = gid; o.inst = -1; o.mat = S.material;
        o.c = from3<R>(S.center); o.r = R(S.radius); o.map = S.mapping;
        V3<R> rv(o.r, o.r, o.r);
        o.box = box_points(o.c - rv, o.c + rv);  // sphere.cuh:50-52
        s.objs.push_back(o);
    }
    for (uint32_t i = 0; i < d->num_quads; ++i, ++gid) {
        const crt_quad& Q = d->quads[i];
        if (Q.material >= d->num_materials) return CRT_ERR_INVALID_SCENE;
        Obj<R> o{};
        o.type = OBJ_QUAD; o.gid = gid; o.inst = -1; o.mat = Q.material;
        o.q = from3<R>(Q.corner); o.qu = from3<R>(Q.u); o.qv = from3<R>(Q.v);
        // quad ctor (core/src/primitives/quad.cuh:66-79)
        V3<R> n = cross(o.qu, o.qv);
        o.qn = unit(n); o.qd = dot(o.qn, o.q); o.qw = n / dot(n, n); o.area = length(n);
        if (ref_style) {
            o.box = pad(box_points(o.q, o.q + o.qu + o.qv));  // quad.cuh:96-99 (diagonal only)
        } else {
            Box<R> b = box_points(o.q, o.q + o.qu);
            b = merge(b, box_points(o.q + o.qv, o.q + o.qu + o.qv));
            o.box = pad(b);
        }
        s.objs.push_back(o);
    }
    // disks, cylinders and cones (created after the quads and boxes, core/src/gpu.cu:565-595); the
    // box holds the corners of the shape's own box (disk.cuh:49, cylinder.cuh:58-61, cone.cuh:52-55)
    // mapped to the world, padded
    if (d->num_shapes && !d->shapes) return CRT_ERR_INVALID_SCENE;
    for (uint32_t i = 0; i < d->num_shapes; ++i, ++gid) {
        const crt_shape& H = d->shapes[i];
        if (H.material >= d->num_materials || H.kind > CRT_SHAPE_CONE) return CRT_ERR_INVALID_SCENE;
        Obj<R> o{};
        o.type = OBJ_SHAPE; o.gid = gid; o.inst = -1; o.mat = H.material;
        o.kind = H.kind; o.c = from3<R>(H.center); o.r = R(H.radius); o.h = R(H.height); o.map = H.mapping;
        o.xf.set(H.transform);
        const R rad = std::fabs(o.r);
        const R ylo = H.kind == CRT_SHAPE_DISK ? -o.h / R(2) : std::min(R(0), o.h);
        const R yhi = H.kind == CRT_SHAPE_DISK ? o.h / R(2) : std::max(R(0), o.h);
        Box<R> b{o.xf.point_to_world(o.c + V3<R>(-rad, ylo, -rad)), o.xf.point_to_world(o.c + V3<R>(-rad, ylo, -rad))};
        for (int k = 1; k < 8; ++k) {
            V3<R> w = o.xf.point_to_world(o.c + V3<R>((k & 1) ? rad : -rad, (k & 2) ? yhi : ylo, (k & 4) ? rad : -rad));
            b = merge(b, box_points(w, w));
        }
        R ext = 0;
        for (int a = 0; a < 3; ++a) ext = std::max(ext, std::max(std::fabs(b.lo[a]), std::fabs(b.hi[a])));
        const V3<R> pd(R(1e-5) * ext + R(1e-6), R(1e-5) * ext + R(1e-6), R(1e-5) * ext + R(1e-6));
        o.box = pad(Box<R>{b.lo - pd, b.hi + pd});
        s.objs.push_back(o);
    }
    auto add_mesh = [&](const crt_mesh& M, const float* xf, int32_t inst, int32_t mat_override, int32_t gid0) -> int {
        for (uint32_t t = 0; t < M.num_triangles; ++t) {
            uint32_t i0 = M.indices[3 * t], i1 = M.indices[3 * t + 1], i2 = M.indices[3 * t + 2];
            if (i0 >= M.num_vertices || i1 >= M.num_vertices || i2 >= M.num_vertices) return CRT_ERR_INVALID_SCENE;
            V3<R> a = from3<R>(M.positions + 3 * i0), b = from3<R>(M.positions + 3 * i1), c = from3<R>(M.positions + 3 * i2);
            if (xf) { a = xform_point(xf, a); b = xform_point(xf, b); c = xform_point(xf, c); }
            uint32_t mat = M.tri_materials ? M.tri_materials[t] : M.material;
            if (mat_override >= 0) mat = (uint32_t)mat_override;
            if (mat >= d->num_materials) return CRT_ERR_INVALID_SCENE;
            const float* na = M.normals ? M.normals + 3 * i0 : nullptr;
            const float* nb = M.normals ? M.normals + 3 * i1 : nullptr;
            const float* nc = M.normals ? M.normals + 3 * i2 : nullptr;
            float nbuf[3][3];
            if (na && xf) {
                // normals of instanced meshes: rotate with the 3x3 part (rigid transforms)
                const float* src[3] = {na, nb, nc};
                for (int k = 0; k < 3; ++k) {
                    V3<R> n = xform_vec(xf, from3<R>(src[k]));
                    nbuf[k][0] = (float)n.x; nbuf[k][1] = (float)n.y; nbuf[k][2] = (float)n.z;
                }
                na = nbuf[0]; nb = nbuf[1]; nc = nbuf[2];
            }
            const float* ta = M.uvs ? M.uvs + 2 * i0 : nullptr;
            const float* tb = M.uvs ? M.uvs + 2 * i1 : nullptr;
            const float* tc = M.uvs ? M.uvs + 2 * i2 : nullptr;
            add_triangle(s, a, b, c, na, nb, nc, ta, tb, tc, mat, gid0 + (int32_t)t, inst);
        }
        return CRT_OK;
    };
    if (d->num_instances == 0) {
        for (uint32_t m = 0; m < d->num_meshes; ++m) {
            int rc = add_mesh(d->meshes[m], nullptr, -1, -1, gid);
            if (rc) return rc;
            gid += (int32_t)d->meshes[m].num_triangles;
        }
    } else {
        // instanced triangles keep their mesh-local id; the instance is reported apart
        for (uint32_t k = 0; k < d->num_instances; ++k) {
            const crt_instance& I = d->instances[k];
            if (I.mesh >= d->num_meshes) return CRT_ERR_INVALID_SCENE;
            int rc = add_mesh(d->meshes[I.mesh], I.object_to_world, (int32_t)k, I.material_override, 0);
            if (rc) return rc;
        }
    }
    for (auto& o : s.objs) {
        if (o.type != OBJ_LIGHT && o.type != OBJ_SPHERE_LIGHT && o.mat >= s.mats.size()) return CRT_ERR_INVALID_SCENE;
    }
    return CRT_OK;
}

// Camera: perspective_camera::initialize (core/src/cameras/perspective_camera.cuh:41-104),
// degrees_to_radians = deg * M_PI / 180 (core/src/utilities/math_utils.cuh:55-58); or
// orthographic_camera::initialize (core/src/cameras/orthographic_camera.cuh:57-118) for an image of
// `width` columns: image height int(width / aspect) (at least 1), viewport 2 * ortho_height high and
// as wide as the image's proportions, its lower-left corner half a viewport from lookfrom.
template <class R> static void setup_camera(Scene<R>& s, const crt_camera& c, uint32_t width) {
    V3<R> lookfrom = from3<R>(c.lookfrom), lookat = from3<R>(c.lookat), vup = from3<R>(c.vup);
    R theta = R(c.vfov_deg) * R(3.14159265358979323846) / R(180.0f);
    R focus = R(c.focus_dist), ratio = R(c.aspect_ratio);
    s.time0 = R(c.time0); s.time1 = R(c.time1);
    s.lens_radius = R(c.aperture) / R(2);
    s.cw = unit(lookfrom - lookat);
    s.cu = unit(cross(vup, s.cw));
    s.cv = cross(s.cw, s.cu);
    R half_height = std::tan(theta / R(2));
    R half_width = ratio * half_height;
    s.origin = lookfrom;
    s.llc = s.origin - (half_width * focus) * s.cu - (half_height * focus) * s.cv - focus * s.cw;
    s.horizontal = (R(2) * half_width * focus) * s.cu;
    s.vertical = (R(2) * half_height * focus) * s.cv;
    s.ortho = c.orthographic != 0;
    if (s.ortho) {
        int ih = (int)((float)width / c.aspect_ratio);
        ih = ih < 1 ? 1 : ih;
        R vh = R(2) * R(c.ortho_height);
        R vw = vh * (R(width) / R(ih));
        s.llc = s.origin - (R(0.5f) * vw) * s.cu - (R(0.5f) * vh) * s.cv;
        s.horizontal = vw * s.cu;
        s.vertical = vh * s.cv;
    }
}

// ------------------------------------------------------------------ reference-style BVH
// bvh_node ctor (core/src/misc/bvh_node.cuh:63-108): random axis, sort by bbox min, split at n/2,
// one-object span stores the object in both children.
template <class R> static int build_ref_bvh(Scene<R>& s, std::vector<int>& ids, int start, int end, Minstd& rng) {
    typename Scene<R>::Node node{};
    int axis = rng.uniform_int(0, 2);
    auto cmp = [&](int a, int b) { return s.objs[a].box.lo[axis] < s.objs[b].box.lo[axis]; };
    int span = end - start;
    int me = (int)s.nodes.size();
    s.nodes.push_back(node);
    int L, Rr; bool ll, lr;
    if (span == 1) {
        L = Rr = ids[start]; ll = lr = true;
    } else if (span == 2) {
        if (cmp(ids[start], ids[start + 1])) { L = ids[start]; Rr = ids[start + 1]; }
        else { L = ids[start + 1]; Rr = ids[start]; }
        ll = lr = true;
    } else {
        if (span <= 64) {
            // the reference's exchange sort (bvh_node.cuh:94-100), exact for small spans
            for (int i = start; i < end; ++i)
                for (int j = i + 1; j < end; ++j)
                    if (!cmp(ids[i], ids[j])) std::swap(ids[i], ids[j]);
        } else {
            // O(n log n) replacement with the same comparator (documented deviation: tie order)
            std::sort(ids.begin() + start, ids.begin() + end, cmp);
        }
        int mid = start + span / 2;
        L = build_ref_bvh(s, ids, start, mid, rng); ll = false;
        Rr = build_ref_bvh(s, ids, mid, end, rng); lr = false;
    }
    auto& n = s.nodes[me];
    n.left = L; n.right = Rr; n.leaf_l = ll; n.leaf_r = lr;
    Box<R> bl = ll ? s.objs[L].box : s.nodes[L].box;
    Box<R> br = lr ? s.objs[Rr].box : s.nodes[Rr].box;
    n.box = merge(bl, br);
    return me;
}

// ------------------------------------------------------------------ intersection
template <class R> struct Ray { V3<R> o, d; R time; };
template <class R> struct Hit {
    R t = 0; V3<R> p, normal; R u = 0, v = 0; bool front = true; int obj = -1;
    R bu = 0, bv = 0;  // triangle barycentrics (weights of v1, v2)
};
struct Counters { uint64_t nodes = 0, prims = 0, rays = 0, retraces = 0, light_tests = 0; };

// aabb::hit (core/src/misc/aabb.cuh:141-166)
template <class R> static bool box_hit(const Box<R>& b, const Ray<R>& r, R tmin, R tmax) {
    for (int a = 0; a < 3; ++a) {
        R invD = R(1) / r.d[a];
        R orig = r.o[a];
        R t0 = (b.lo[a] - orig) * invD;
        R t1 = (b.hi[a] - orig) * invD;
        if (invD < 0) std::swap(t0, t1);
        if (t0 > tmin) tmin = t0;
        if (t1 < tmax) tmax = t1;
        if (tmax <= tmin) return false;
    }
    return true;
}

// set_face_normal (core/src/misc/hit_record.cuh:34-38)
template <class R> static void set_face_normal(Hit<R>& h, const Ray<R>& r, V3<R> outward) {
    h.front = dot(r.d, outward) < R(0);
    h.normal = h.front ? outward : -outward;
}

// get_sphere_uv (core/src/utilities/uvmapping.cuh:204-228). M_PI is the float literal of
// core/src/misc/constants.cuh:35, so the reference evaluates this in float (R = float here); the FP64
// oracle uses pi in double.
template <class R> static void sphere_uv(V3<R> p, const crt_uvmap& m, R& u, R& v) {
    R theta = std::acos(-p.y);
    R phi = std::atan2(-p.z, p.x) + R(3.14159265358979323846);
    R s = phi / (R(2) * R(3.14159265358979323846));
    R t = theta / R(3.14159265358979323846);
    s = std::fmod(s * R(m.repeat_u), R(1));
    t = std::fmod(t * R(m.repeat_v), R(1));
    u = R(m.scale_u) * s + R(m.offset_u);
    v = R(m.scale_v) * t + R(m.offset_v);
}

// get_cylinder_uv / get_disk_uv / get_cone_uv (core/src/utilities/uvmapping.cuh:285-368; M_PI is the
// float literal of core/src/misc/constants.cuh:35). The cylinder and the cone are given their
// outward vector, the disk its hit point relative to the centre; the cone applies the u repeat twice.
template <class R> static void finish_uv(R s_, R t_, const crt_uvmap& m, R& u, R& v) {
    u = R(m.scale_u) * s_ + R(m.offset_u);
    v = R(m.scale_v) * t_ + R(m.offset_v);
}
template <class R> static void cylinder_uv(V3<R> p, R h, const crt_uvmap& m, R& u, R& v) {
    const R pi = R(3.14159265358979323846);
    R theta = std::atan2(p.x, p.z);
    R s_ = R(1) - (theta + pi) / (R(2) * pi);
    R t_ = (p.y + h / R(2)) / h;
    finish_uv(std::fmod(s_ * R(m.repeat_u), R(1)), std::fmod(t_ * R(m.repeat_v), R(1)), m, u, v);
}
template <class R> static void disk_uv(V3<R> p, R r, const crt_uvmap& m, R& u, R& v) {
    R s_ = p.x / (R(2) * r) + R(0.5f);
    R t_ = p.z / (R(2) * r) + R(0.5f);
    finish_uv(std::fmod(s_ * R(m.repeat_u), R(1)), std::fmod(t_ * R(m.repeat_v), R(1)), m, u, v);
}
template <class R> static void cone_uv(V3<R> p, R r, const crt_uvmap& m, R& u, R& v) {
    const R pi = R(3.14159265358979323846);
    R theta = std::atan2(p.x, p.z);
    R s_ = (theta + pi) / (R(2) * pi);
    R t_ = std::sqrt(p.x * p.x + p.z * p.z) / r;
    s_ = std::fmod(s_ * R(m.repeat_u), R(1));
    s_ = std::fmod(s_ * R(m.repeat_u), R(1));
    finish_uv(s_, std::fmod(t_ * R(m.repeat_v), R(1)), m, u, v);
}

// get_triangle_uv (core/src/utilities/uvmapping.cuh:415-447, get_barycenter math_utils.cuh:73-86)
template <class R> static void triangle_uv_at(const Obj<R>& o, V3<R> p, R& u, R& v) {
    V3<R> v0 = o.v[1] - o.v[0], v1 = o.v[2] - o.v[0], v2 = p - o.v[0];
    R d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(v2, v0), d21 = dot(v2, v1);
    R denom = d00 * d11 - d01 * d01;
    R bv = (d11 * d20 - d01 * d21) / denom;
    R bw = (d00 * d21 - d01 * d20) / denom;
    R bu = R(1) - bv - bw;
    R uu = bu * R(o.uv[0][0]) + bv * R(o.uv[1][0]) + (R(1) - bu - bv) * R(o.uv[2][0]);
    R vv = bu * R(o.uv[0][1]) + bv * R(o.uv[1][1]) + (R(1) - bu - bv) * R(o.uv[2][1]);
    uu = std::fmod(uu, R(1)); vv = std::fmod(vv, R(1));
    if (uu < 0) uu += R(1);
    if (vv < 0) vv += R(1);
    u = uu; v = vv;
}

// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013), two-sided, inclusive t range.
// Returns t and the barycentric weights of v1 (bu) and v2 (bv).
template <class R> static bool watertight(const V3<R>& A0, const V3<R>& B0, const V3<R>& C0, const Ray<R>& r,
                                          R tmin, R tmax, R& t, R& bu, R& bv) {
    R ax = std::fabs(r.d.x), ay = std::fabs(r.d.y), az = std::fabs(r.d.z);
    int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (r.d[kz] < 0) std::swap(kx, ky);
    R Sx = r.d[kx] / r.d[kz], Sy = r.d[ky] / r.d[kz], Sz = R(1) / r.d[kz];
    V3<R> A = A0 - r.o, B = B0 - r.o, C = C0 - r.o;
    R Ax = A[kx] - Sx * A[kz], Ay = A[ky] - Sy * A[kz];
    R Bx = B[kx] - Sx * B[kz], By = B[ky] - Sy * B[kz];
    R Cx = C[kx] - Sx * C[kz], Cy = C[ky] - Sy * C[kz];
    R U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if (U == R(0) || V == R(0) || W == R(0)) {
        double CxBy = (double)Cx * By, CyBx = (double)Cy * Bx; U = (R)(CxBy - CyBx);
        double AxCy = (double)Ax * Cy, AyCx = (double)Ay * Cx; V = (R)(AxCy - AyCx);
        double BxAy = (double)Bx * Ay, ByAx = (double)By * Ax; W = (R)(BxAy - ByAx);
    }
    if ((U < 0 || V < 0 || W < 0) && (U > 0 || V > 0 || W > 0)) return false;
    R det = U + V + W;
    if (det == R(0)) return false;
    R Az = Sz * A[kz], Bz = Sz * B[kz], Cz = Sz * C[kz];
    R T = U * Az + V * Bz + W * Cz;
    R inv = R(1) / det;
    R tt = T * inv;
    if (!(tt >= tmin && tt <= tmax)) return false;
    t = tt; bu = V * inv; bv = W * inv;
    return true;
}

// One object test with the reference's interval convention per type.
template <class R> static bool obj_hit(const Scene<R>& s, int idx, const Ray<R>& r, R tmin, R tmax, bool camera,
                                       Hit<R>& rec) {
    const Obj<R>& o = s.objs[idx];
    if (o.type == OBJ_SPHERE || o.type == OBJ_SPHERE_LIGHT) {
        // sphere::hit (core/src/primitives/sphere.cuh:111-166): strict interval (surrounds); the
        // sphere lights repeat it and then hide from camera rays when invisible (omni_light.cuh:40-78)
        V3<R> oc = r.o - o.c;
        R a = len2(r.d);
        R half_b = dot(oc, r.d);
        R c = len2(oc) - o.r * o.r;
        R root;
        if (s.ref_style) {
            R disc = half_b * half_b - a * c;
            if (disc < 0) return false;
            R sq = std::sqrt(disc);
            root = (-half_b - sq) / a;
            if (!(tmin < root && root < tmax)) {
                root = (-half_b + sq) / a;
                if (!(tmin < root && root < tmax)) return false;
            }
        } else {
            // robust form (perpendicular-distance discriminant, stable root pair)
            V3<R> dn = r.d / std::sqrt(a);
            V3<R> l = oc - dot(oc, dn) * dn;
            R disc = a * (o.r * o.r - len2(l));
            if (disc < 0) return false;
            R sq = std::sqrt(disc);
            R q = half_b >= 0 ? -(half_b + sq) : -(half_b - sq);
            R r0 = c / q, r1 = q / a;
            if (q == R(0)) { r0 = r1 = -half_b / a; }
            if (r0 > r1) std::swap(r0, r1);
            root = r0;
            if (!(tmin < root && root < tmax)) {
                root = r1;
                if (!(tmin < root && root < tmax)) return false;
            }
        }
        if (o.type == OBJ_SPHERE_LIGHT && o.invisible && camera) return false;
        rec.t = root;
        rec.p = r.o + root * r.d;
        V3<R> outward = (rec.p - o.c) / o.r;
        set_face_normal(rec, r, outward);
        sphere_uv(outward, o.map, rec.u, rec.v);
        rec.obj = idx;
        return true;
    }
    if (o.type == OBJ_SHAPE) {
        // the ray in the shape's space (its wrappers), then disk::hit (core/src/primitives/disk.cuh:
        // 53-82), cylinder::hit (cylinder.cuh:63-103) or cone::hit (cone.cuh:58-128) on the closed
        // interval. Written so that a NaN distance (a ray parallel to the disk, or along the
        // cylinder's axis) is no hit; the reference's comparisons let some of those through.
        V3<R> ro = r.o, rd = r.d;
        o.xf.to_object(ro, rd);
        R t;
        V3<R> outward, uvp;
        if (o.kind == CRT_SHAPE_DISK) {
            t = (o.c.y - ro.y) / rd.y;
            if (!(tmin <= t && t <= tmax)) return false;
            V3<R> hp = ro + t * rd;
            R dist2 = (hp.x - o.c.x) * (hp.x - o.c.x) + (hp.z - o.c.z) * (hp.z - o.c.z);
            if (!(dist2 <= o.r * o.r)) return false;
            outward = V3<R>(0, 1, 0);
            uvp = hp - o.c;
        } else {
            R a, b, c, disc;
            if (o.kind == CRT_SHAPE_CYLINDER) {
                V3<R> oc = ro - o.c;
                a = rd.x * rd.x + rd.z * rd.z;
                b = R(2) * (oc.x * rd.x + oc.z * rd.z);
                c = oc.x * oc.x + oc.z * oc.z - o.r * o.r;
                disc = b * b - R(4) * a * c;
                if (!(disc >= 0)) return false;
            } else {
                R A = ro.x - o.c.x, B = ro.z - o.c.z, D = o.h - ro.y + o.c.y;
                R tn = (o.r / o.h) * (o.r / o.h);
                a = rd.x * rd.x + rd.z * rd.z - tn * (rd.y * rd.y);
                b = R(2) * A * rd.x + R(2) * B * rd.z + R(2) * tn * D * rd.y;
                c = A * A + B * B - tn * (D * D);
                disc = b * b - R(4) * (a * c);
                if (std::fabs(disc) < R(0.001f) || !(disc >= 0)) return false;
            }
            R sq = std::sqrt(disc);
            t = (-b - sq) / (R(2) * a);
            if (!(tmin <= t && t <= tmax)) {
                t = (-b + sq) / (R(2) * a);
                if (!(tmin <= t && t <= tmax)) return false;
            }
            R y = ro.y + t * rd.y;
            if (!(y >= o.c.y && y <= o.c.y + o.h)) return false;
            V3<R> hp = ro + t * rd;
            if (o.kind == CRT_SHAPE_CYLINDER) {
                outward = (hp - o.c) / o.r;   // with its y part, not made unit (cylinder.cuh:93-95)
                uvp = outward;
            } else {
                if (hp.y <= o.c.y) uvp = V3<R>(0, -1, 0);
                else if (hp.y >= o.c.y + o.h) uvp = V3<R>(0, 1, 0);
                else {
                    R rs = std::sqrt((hp.x - o.c.x) * (hp.x - o.c.x) + (hp.z - o.c.z) * (hp.z - o.c.z));
                    uvp = V3<R>(hp.x - o.c.x, rs * (o.r / o.h), hp.z - o.c.z);
                }
                outward = unit(uvp);
            }
        }
        // set_face_normal in the shape's space, then the normal back through the wrappers
        rec.front = dot(rd, outward) < R(0);
        rec.normal = o.xf.normal_to_world(rec.front ? outward : -outward);
        if (o.kind == CRT_SHAPE_DISK) disk_uv(uvp, o.r, o.map, rec.u, rec.v);
        else if (o.kind == CRT_SHAPE_CYLINDER) cylinder_uv(uvp, o.h, o.map, rec.u, rec.v);
        else cone_uv(uvp, o.r, o.map, rec.u, rec.v);
        rec.t = t;
        rec.p = r.o + t * r.d;
        rec.obj = idx;
        return true;
    }
    if (o.type == OBJ_QUAD || o.type == OBJ_LIGHT) {
        // quad::hit (core/src/primitives/quad.cuh:100-133), directional_light::hit (:76-118)
        R denom = dot(o.qn, r.d);
        if (std::fabs(denom) < R(1e-8)) return false;
        R t = (o.qd - dot(o.qn, r.o)) / denom;
        if (!(tmin <= t && t <= tmax)) return false;
        V3<R> p = r.o + t * r.d;
        V3<R> hp = p - o.q;
        R alpha = dot(o.qw, cross(hp, o.qv));
        R beta = dot(o.qw, cross(o.qu, hp));
        if (o.type == OBJ_QUAD) {
            if (alpha < 0 || R(1) < alpha || beta < 0 || R(1) < beta) return false;
            rec.u = alpha; rec.v = beta;
        } else {
            if (alpha < R(-0.5) || R(0.5) < alpha || beta < R(-0.5) || R(0.5) < beta) return false;
            rec.u = alpha + R(0.5f); rec.v = beta + R(0.5f);   // written before the visibility test
            if (o.invisible && camera) return false;
        }
        rec.t = t; rec.p = p;
        set_face_normal(rec, r, o.qn);
        rec.obj = idx;
        return true;
    }
    // triangle
    R t, bu, bv;
    if (s.ref_style) {
        // Moller-Trumbore (core/src/primitives/triangle.cuh:130-185), |det| < 1e-6 rejected
        V3<R> pv = cross(r.d, o.e2);
        R det = dot(o.e1, pv);
        if (std::fabs(det) < R(0.000001f)) return false;
        R inv = R(1.0f) / det;
        V3<R> tv = r.o - o.v[0];
        R u = dot(tv, pv) * inv;
        if (u < 0 || u > 1) return false;
        V3<R> qv = cross(tv, o.e1);
        R v = dot(r.d, qv) * inv;
        if (v < 0 || u + v > 1) return false;
        t = dot(o.e2, qv) * inv;
        if (t < tmin || t > tmax) return false;
        bu = u; bv = v;
        // UV from the stale rec.hit_point (triangle.cuh:162 before :164)
        if (o.has_uv) triangle_uv_at(o, rec.p, rec.u, rec.v); else { rec.u = 0; rec.v = 0; }
    } else {
        if (!watertight(o.v[0], o.v[1], o.v[2], r, tmin, tmax, t, bu, bv)) return false;
    }
    rec.t = t;
    rec.p = r.o + t * r.d;
    rec.bu = bu; rec.bv = bv;
    if (!s.ref_style) {
        if (o.has_uv) triangle_uv_at(o, rec.p, rec.u, rec.v); else { rec.u = 0; rec.v = 0; }
    }
    V3<R> n = o.middle_normal;
    if (o.smooth) n = bu * o.n[1] + bv * o.n[2] + (R(1) - bu - bv) * o.n[0];  // triangle.cuh:167-174
    set_face_normal(rec, r, n);
    rec.obj = idx;
    return true;
}

// bvh_node::hit (core/src/misc/bvh_node.cuh:110-118): left then right with the shrunk interval.
template <class R> static bool bvh_hit(const Scene<R>& s, int node, bool leaf, const Ray<R>& r, R tmin, R tmax,
                                       bool camera, Hit<R>& rec, Counters& cnt) {
    if (leaf) { cnt.prims++; return obj_hit(s, node, r, tmin, tmax, camera, rec); }
    const auto& n = s.nodes[node];
    cnt.nodes++;
    if (!box_hit(n.box, r, tmin, tmax)) return false;
    bool hl = bvh_hit(s, n.left, n.leaf_l, r, tmin, tmax, camera, rec, cnt);
    bool hr = bvh_hit(s, n.right, n.leaf_r, r, tmin, hl ? rec.t : tmax, camera, rec, cnt);
    return hl || hr;
}

// The build's deterministic closest hit (oracle64 only; refcpu keeps the reference's traversal
// order): hits within a relative 2^-20 of each other in t tie and the lower global primitive id
// wins. A candidate is tested up to bt * (1 + 2^-20) and taken when it is nearer than
// bt * (1 - 2^-20) or ties with a lower id; with no hit yet the id to beat is 0, so a first hit
// must lie below tmax by the same margin. Both factors are exact in FP32 and FP64.
constexpr double kTieHi = 1.0 + 1.0 / 1048576.0, kTieLo = 1.0 - 1.0 / 1048576.0;
template <class R> static inline bool tie_wins(R t, uint32_t key, R bt, uint32_t bkey) {
    return t < bt * R(kTieLo) || key < bkey;
}
template <class R> static void bvh_hit_tie(const Scene<R>& s, int node, bool leaf, const Ray<R>& r, R tmin, bool camera,
                                           R& bt, uint32_t& bkey, Hit<R>& rec, bool& any, Counters& cnt) {
    if (leaf) {
        cnt.prims++;
        Hit<R> tmp;
        const uint32_t key = (uint32_t)s.objs[node].gid;
        if (obj_hit(s, node, r, tmin, bt * R(kTieHi), camera, tmp) && tie_wins(tmp.t, key, bt, bkey)) {
            rec = tmp; bt = tmp.t; bkey = key; any = true;
        }
        return;
    }
    const auto& n = s.nodes[node];
    cnt.nodes++;
    if (!box_hit(n.box, r, tmin, bt * R(kTieHi))) return;
    bvh_hit_tie(s, n.left, n.leaf_l, r, tmin, camera, bt, bkey, rec, any, cnt);
    bvh_hit_tie(s, n.right, n.leaf_r, r, tmin, camera, bt, bkey, rec, any, cnt);
}

// hittable_list::hit over the world (core/src/primitives/hittable_list.cuh:149-166).
template <class R> static bool world_hit(const Scene<R>& s, const Ray<R>& r, R tmin, R tmax, bool camera, Hit<R>& rec,
                                         Counters& cnt) {
    if (!s.ref_style) {
        R bt = tmax;
        uint32_t bkey = 0;
        bool any = false;
        if (s.accel == 1 || s.root < 0) {
            for (size_t i = 0; i < s.objs.size(); ++i) {
                cnt.prims++;
                Hit<R> tmp;
                const uint32_t key = (uint32_t)s.objs[i].gid;
                if (obj_hit(s, (int)i, r, tmin, bt * R(kTieHi), camera, tmp) && tie_wins(tmp.t, key, bt, bkey)) {
                    rec = tmp; bt = tmp.t; bkey = key; any = true;
                }
            }
        } else {
            bvh_hit_tie(s, s.root, s.root_is_leaf, r, tmin, camera, bt, bkey, rec, any, cnt);
        }
        return any;
    }
    Hit<R> tmp;  // fresh record: hit_point starts at (0,0,0)
    bool any = false;
    if (s.accel == 1 || s.root < 0) {
        R closest = tmax;
        for (size_t i = 0; i < s.objs.size(); ++i) {
            cnt.prims++;
            if (obj_hit(s, (int)i, r, tmin, closest, camera, tmp)) { any = true; closest = tmp.t; rec = tmp; }
        }
        return any;
    }
    if (bvh_hit(s, s.root, s.root_is_leaf, r, tmin, tmax, camera, tmp, cnt)) { rec = tmp; return true; }
    return false;
}

// ------------------------------------------------------------------ textures
template <class R> static V3<R> tex_value(const Scene<R>& s, int ti, R u, R v, V3<R> p, int depth = 0) {
    const crt_texture& t = s.texs[ti];
    if (t.type == CRT_TEX_SOLID) return from3<R>(t.color);
    if (t.type == CRT_TEX_CHECKER) {
        // checker_texture::value (core/src/textures/checker_texture.cuh:33-42)
        R inv = R(1.0f) / R(t.scale);
        int xi = (int)std::floor(inv * p.x), yi = (int)std::floor(inv * p.y), zi = (int)std::floor(inv * p.z);
        bool even = (xi + yi + zi) % 2 == 0;
        int sub = even ? t.even : t.odd;
        if (depth > 0 || sub < 0 || sub >= (int)s.texs.size()) return V3<R>();
        return tex_value(s, sub, u, v, p, depth + 1);
    }
    // image_texture::value (core/src/textures/image_texture.cuh:54-71) + bitmap_image::pixel_data
    // (core/src/utilities/bitmap_image.cuh:74-91): clamp uv, flip v, nearest texel, clamp index.
    if (t.height == 0 || s.tex_pixels[ti].empty()) return V3<R>(R(0), R(1), R(1));
    R uu = u < 0 ? R(0) : (u > 1 ? R(1) : u);
    R vv = v < 0 ? R(0) : (v > 1 ? R(1) : v);
    vv = R(1.0f) - vv;
    int i = (int)(uu * R(t.width)), j = (int)(vv * R(t.height));
    auto clampi = [](int x, int lo, int hi) { return x < lo ? lo : (x < hi ? x : hi - 1); };
    i = clampi(i, 0, (int)t.width); j = clampi(j, 0, (int)t.height);
    const uint8_t* px = s.tex_pixels[ti].data() + ((size_t)j * t.width + i) * t.channels;
    R sc = R(1.0f) / R(255.0f);
    return {sc * R(px[0]), sc * R(px[1]), sc * R(px[2])};
}

template <class R> static V3<R> albedo_of(const Scene<R>& s, const crt_material& m, const Hit<R>& h) {
    if (m.texture >= 0 && m.texture < (int)s.texs.size()) return tex_value(s, m.texture, h.u, h.v, h.p);
    return from3<R>(m.albedo);
}

// ------------------------------------------------------------------ sampling helpers
// random_in_unit_disk (core/src/misc/gpu_randomizer.cuh:64-71); x drawn before y.
template <class R> static V3<R> random_in_unit_disk(Rng& g) {
    V3<R> p;
    do {
        R a = R(g.uniform()); R b = R(g.uniform());
        p = V3<R>(R(2.0f) * a - R(1), R(2.0f) * b - R(1), R(0));
    } while (dot(p, p) >= R(1.0f));
    return p;
}
// random_in_unit_sphere (gpu_randomizer.cuh:73-80); l1() is the squared length.
template <class R> static V3<R> random_in_unit_sphere(Rng& g) {
    V3<R> p;
    do {
        R a = R(g.uniform()); R b = R(g.uniform()); R c = R(g.uniform());
        p = V3<R>(R(2.0f) * a - R(1), R(2.0f) * b - R(1), R(2.0f) * c - R(1));
    } while (len2(p) >= R(1.0f));
    return p;
}
// random_cosine_direction (gpu_randomizer.cuh:122-133)
template <class R> static V3<R> random_cosine_direction(Rng& g) {
    R r1 = R(g.uniform()); R r2 = R(g.uniform());
    R phi = R(6.28318530718f) * r1;
    R z = std::sqrt(R(1) - r2);
    R rs = std::sqrt(r2);
    return {std::cos(phi) * rs, std::sin(phi) * rs, z};
}
// onb::build_from_w (core/src/misc/onb.cuh:122-128)
template <class R> struct Onb {
    V3<R> u, v, w;
    explicit Onb(V3<R> n) {
        w = unit(n);
        V3<R> a = (std::fabs(w.x) > R(0.9f)) ? V3<R>(0, 1, 0) : V3<R>(1, 0, 0);
        v = unit(cross(w, a));
        u = cross(w, v);
    }
    V3<R> local(V3<R> a) const { return a.x * u + a.y * v + a.z * w; }
};

// directional_light::pdf_value (core/src/lights/directional_light.cuh:147-158); the hit uses
// depth 0 so the light is always visible here. Sphere lights (omni_light.cuh:80-88,
// spot_light.cuh:96-104): one over the solid angle of the sphere's cone seen from o.
template <class R> static R light_pdf_value(const Scene<R>& s, int li, V3<R> o, V3<R> v, Counters& cnt) {
    Hit<R> rec;
    Ray<R> r{o, v, 0};
    cnt.light_tests++;
    if (!obj_hit(s, li, r, R(0.00001f), std::numeric_limits<R>::infinity(), false, rec)) return 0;
    const Obj<R>& L = s.objs[li];
    if (L.type == OBJ_SPHERE_LIGHT) {
        R cos_max = std::sqrt(R(1) - L.r * L.r / len2(L.c - o));
        R solid = R(2) * R(3.14159265358979323846) * (R(1) - cos_max);
        return R(1) / solid;
    }
    R dist2 = rec.t * rec.t * len2(v);
    R cosine = std::fabs(dot(v, rec.normal) / length(v));
    return dist2 / (cosine * s.objs[li].area);
}
// hittable_list::pdf_value (hittable_list.cuh:173-185)
template <class R> static R lights_pdf_value(const Scene<R>& s, V3<R> o, V3<R> v, Counters& cnt) {
    R weight = R(1.0f) / R((float)s.lights.size());
    R sum = 0;
    for (size_t i = 0; i < s.lights.size(); ++i) sum += weight * light_pdf_value(s, s.lights[i], o, v, cnt);
    return sum;
}
// hittable_list::random (hittable_list.cuh:191-216) -> directional_light::random (:165-172), or
// for a sphere light omni_light::random (omni_light.cuh:90-98): random_to_sphere
// (gpu_randomizer.cuh:82-93) in the frame built on the direction to the centre.
template <class R> static V3<R> lights_random(const Scene<R>& s, V3<R> o, Rng& g) {
    if (s.lights.empty()) return V3<R>();
    int r = 0;
    if (s.lights.size() > 1) r = g.uniform_int_min(0, (int)s.lights.size() - 1);
    const Obj<R>& L = s.objs[s.lights[r]];
    if (L.type == OBJ_SPHERE_LIGHT) {
        V3<R> dir = L.c - o;
        R d2 = len2(dir);
        Onb<R> uvw(dir);
        R r1 = R(g.uniform());
        R r2 = R(g.uniform());
        R z = R(1) + r2 * (std::sqrt(R(1) - L.r * L.r / d2) - R(1));
        R phi = R(2) * R(3.14159265358979323846) * r1;
        R x = std::cos(phi) * std::sqrt(R(1) - z * z);
        R y = std::sin(phi) * std::sqrt(R(1) - z * z);
        return uvw.local(V3<R>(x, y, z));
    }
    R a = R(g.uniform()); R b = R(g.uniform());
    V3<R> p = L.q + (a - R(0.5f)) * L.qu + (b - R(0.5f)) * L.qv;
    return p - o;
}

// ------------------------------------------------------------------ scatter directions
// metal::scatter (core/src/materials/metal.cuh:57-65): mirror of the unit incoming direction about
// the normal (glm::reflect) plus fuzz times a point of the unit ball; never absorbs.
template <class R> static V3<R> metal_dir(V3<R> d, V3<R> n, R fuzz, Rng& g) {
    V3<R> refl = reflect(unit(d), n);
    return refl + fuzz * random_in_unit_sphere<R>(g);
}

// dielectric::scatter (core/src/materials/dielectric.cuh:32-51) with reflectance (Schlick on the
// refraction ratio, :57-63). `draw` stands for get_real(rng): it is consumed only when refraction
// is possible (the || short-circuits on total internal reflection).
template <class R, class Draw> static V3<R> dielectric_dir(V3<R> d, V3<R> n, bool front, R ir, Draw draw,
                                                         bool* reflected) {
    R ratio = front ? (R(1.0f) / ir) : ir;
    V3<R> ud = unit(d);
    R dn = dot(-ud, n);
    R cos_t = dn < R(1.0f) ? dn : R(1.0f);  // ffmin (core/src/misc/color.cuh:222-223)
    R sin_t = std::sqrt(R(1.0f) - cos_t * cos_t);
    bool cannot = ratio * sin_t > R(1.0);
    bool do_reflect = cannot;
    if (!cannot) {
        R r0 = (R(1) - ratio) / (R(1) + ratio);
        r0 = r0 * r0;
        R refl = r0 + (R(1) - r0) * std::pow(R(1) - cos_t, R(5));
        do_reflect = refl > R(draw());
    }
    if (reflected) *reflected = do_reflect;
    return do_reflect ? reflect(ud, n) : refract(ud, n, ratio);
}

// Camera ray of image-plane coordinates (u, v): perspective_camera::get_ray
// (core/src/cameras/perspective_camera.cuh:106-113): lens offset from random_in_unit_disk scaled by
// the lens radius in the camera's (u, v) plane, direction to the focus-plane point, time draw.
// The orthographic camera's get_ray (orthographic_camera.cuh:125-142) draws the same way and
// starts the ray at the viewport point plus the offset, along -w.
template <class R> static Ray<R> camera_ray(const Scene<R>& s, R u, R v, Rng& g) {
    V3<R> rdisk = s.lens_radius * random_in_unit_disk<R>(g);
    V3<R> off = s.cu * rdisk.x + s.cv * rdisk.y;
    R tm = s.time0 + R(g.uniform()) * (s.time1 - s.time0);
    if (s.ortho) return Ray<R>{s.llc + u * s.horizontal + v * s.vertical + off, R(-1) * s.cw, tm};
    return Ray<R>{s.origin + off, s.llc + u * s.horizontal + v * s.vertical - s.origin - off, tm};
}

// ------------------------------------------------------------------ estimator
// camera::ray_color, iterative version (core/src/cameras/camera.cuh:149-264).
template <class R> static V3<R> ray_color(const Scene<R>& s, Ray<R> ray, int max_depth, Rng& g, Counters& cnt) {
    V3<R> result(0, 0, 0);
    V3<R> att(1, 1, 1);
    const R INF = std::numeric_limits<R>::max();   // FLT_MAX
    uint32_t bounce = 0;
    for (int depth = max_depth; depth > 0; --depth) {
        Hit<R> rec;
        bool camera = depth == max_depth;
        cnt.rays++;
        if (!world_hit(s, ray, R(0.00001f), INF, camera, rec, cnt)) {
            result = result + att * s.background;
            break;
        }
        ++bounce;
        g.bounce(bounce);
        const Obj<R>* o = &s.objs[rec.obj];
        // emitted (core/src/materials/diffuse_light.cuh:45-67; base material: black, alpha 1)
        V3<R> emitted(0, 0, 0);
        bool alpha0 = false;
        if (o->type == OBJ_LIGHT) {
            if (rec.front) emitted = o->c * o->r; else alpha0 = true;
        } else if (o->type == OBJ_SPHERE_LIGHT) {
            if (!o->spot) {
                emitted = o->le * o->lint;   // diffuse_light, non-directional
            } else if (!rec.front) {
                // diffuse_spot_light::emitted (core/src/materials/diffuse_spot_light.cuh:38-61): back
                // faces give colour(0,0,0,0) when invisible, colour(1,1,1,0) when visible
                alpha0 = true;
                if (!o->invisible) emitted = V3<R>(1, 1, 1);
            } else {
                R ct = dot(unit(ray.d), o->sdir);
                emitted = ct < o->scut ? V3<R>(1, 1, 1) : o->le * (o->lint * std::pow(ct, o->sfall));
            }
        } else if (s.mats[o->mat].type == CRT_MAT_DIFFUSE_LIGHT) {
            const crt_material& m = s.mats[o->mat];
            bool one_sided = (m.flags & CRT_MATF_ONE_SIDED) != 0;
            if (!one_sided || rec.front) emitted = albedo_of(s, m, rec) * R(m.intensity); else alpha0 = true;
        }
        if (alpha0) {
            // pass-through re-trace from t + 0.001 (camera.cuh:179-183); a miss keeps the old record
            cnt.retraces++; cnt.rays++;
            Hit<R> rec2;
            if (world_hit(s, ray, rec.t + R(0.001f), INF, camera, rec2, cnt)) rec = rec2;
            o = &s.objs[rec.obj];
        }
        // scatter (lambertian.cuh:63-136, metal.cuh:57-65, dielectric.cuh:32-51); lights do not
        bool is_light = o->type == OBJ_LIGHT || o->type == OBJ_SPHERE_LIGHT || s.mats[o->mat].type == CRT_MAT_DIFFUSE_LIGHT;
        if (is_light) { result = result + att * emitted; break; }
        const crt_material& m = s.mats[o->mat];
        result = result + att * emitted;
        if (m.type == CRT_MAT_METAL) {
            V3<R> alb = albedo_of(s, m, rec);
            ray = Ray<R>{rec.p, metal_dir(ray.d, rec.normal, R(m.fuzz), g), ray.time};
            att = att * alb;
            continue;
        }
        if (m.type == CRT_MAT_DIELECTRIC) {
            V3<R> dir = dielectric_dir(ray.d, rec.normal, rec.front, R(m.ior), [&]() { return g.uniform(); }, nullptr);
            ray = Ray<R>{rec.p, dir, ray.time};
            continue;  // attenuation 1
        }
        const bool iso = m.type == CRT_MAT_ISOTROPIC, on = m.type == CRT_MAT_OREN_NAYAR;
        if (m.type != CRT_MAT_LAMBERTIAN && !iso && !on) break;  // unsupported material: no scatter
        if (on && s.lights.empty()) break;   // the reference reads lights.objects[0] regardless
        V3<R> alb = albedo_of(s, m, rec);
        if (on) {
            // oren_nayar::scatter (core/src/materials/oren_nayar.cuh): random_on_hemisphere's
            // get_unit_vector draws two values for a direction it never uses; the attenuation is
            // light 0's colour * intensity * albedo * albedo_temp * (1 / pi) * (A + B alpha beta)
            // with sigma = roughness^2, A = 1 - 0.5 sigma / (sigma + 0.33), B = 0.45 sigma /
            // (sigma + 0.09), and the outgoing cosine of its unset skip_pdf_ray, 0 (sin = 1)
            g.uniform(); g.uniform();
            const Obj<R>& l0 = s.objs[s.lights[0]];
            V3<R> inc = l0.type == OBJ_LIGHT ? l0.c * l0.r : l0.le * l0.lint;
            R sg = R(m.fuzz) * R(m.fuzz);
            R A = R(1) - R(0.5f) * (sg / (sg + R(0.33f))), B = R(0.45f) * (sg / (sg + R(0.09f)));
            R cosi = dot(-ray.d, rec.normal);
            R mc = std::fmax(R(0), cosi);
            R si = std::sqrt(R(1) - mc * mc);
            R alpha = std::fmax(si, R(1)) / std::fmin(si, R(1)), beta = std::fmin(si, R(1)) / std::fmax(si, R(1));
            R dr = A + B * alpha * beta;
            alb = inc * alb * R(m.ior) * (R(1) / R(3.14159265358979323846)) * dr;
        }
        const bool sph = iso || on;
        // mixture_pdf(hittable_pdf(lights, p), cosine_pdf(n)) (camera.cuh:203-216, mixture_pdf.cuh:28-48);
        // isotropic (isotropic.cuh) puts sphere_pdf in the cosine's place (sphere_pdf.cuh:20-28): a
        // uniform direction from get_unit_vector (gpu_randomizer.cuh:53-60: angle u * 2pi, then
        // z = 2u - 1, get_real(lo, hi) = u (hi - lo) + lo), pdf and scattering pdf 1 / (4 pi)
        Onb<R> uvw(rec.normal);
        V3<R> dir;
        if (g.uniform() < 0.5f) dir = lights_random(s, rec.p, g);
        else if (sph) {
            R a = R(g.uniform()) * R(6.28318530718f);
            R z = R(g.uniform()) * R(2) - R(1);
            R r = std::sqrt(R(1) - z * z);
            dir = V3<R>(r * std::cos(a), r * std::sin(a), z);
        } else dir = uvw.local(random_cosine_direction<R>(g));
        R pl = lights_pdf_value(s, rec.p, dir, cnt);
        const R inv4pi = R(1) / (R(4) * R(3.14159265358979323846));
        R cv = dot(unit(dir), uvw.w) / R(3.14159265358979323846);
        R pc = sph ? inv4pi : (R(0) > cv ? R(0) : cv);  // ffmax(0, x)
        R pdf = R(0.5f) * pl + (R(1.0f) - R(0.5f)) * pc;
        R ct = dot(rec.normal, unit(dir));
        R spdf = iso ? inv4pi : (ct < 0 ? R(0) : ct / R(3.14159265358979323846));
        // colour / float is (1/t) * colour (core/src/misc/color.cuh:105-108)
        att = att * ((R(1) / pdf) * (alb * spdf));
        ray = Ray<R>{rec.p, dir, ray.time};
    }
    return result;
}

struct RenderOpts {
    int precision = 32;
    int accel = 0;
    int threads = 1;
    uint32_t row_begin = 0, row_end = 0;
    uint32_t pixel_stride = 1;   // render every k-th pixel of a row (sampling the workload)
    bool ref_style = true;
};

template <class R> static int render_impl(const crt_scene_desc* d, const crt_render_desc* rd, const RenderOpts& o,
                                          float* accum, Counters& total, double& secs) {
    Scene<R> s;
    int rc = build_scene(s, d, o.ref_style, o.accel);
    if (rc) return rc;
    setup_camera(s, d->camera, rd->width);
    if (o.accel == 0 && !s.objs.empty()) {
        std::vector<int> ids(s.objs.size());
        for (size_t i = 0; i < ids.size(); ++i) ids[i] = (int)i;
        Minstd brng(1984);  // load_scene seed (core/src/gpu.cu:2068, bvh at :749)
        s.nodes.reserve(2 * ids.size());
        s.root = build_ref_bvh(s, ids, 0, (int)ids.size(), brng);
        s.root_is_leaf = false;
    }
    const uint32_t W = rd->width, H = rd->height;
    int sq = (int)std::sqrt((double)rd->spp);
    uint32_t rb = o.row_begin, re = o.row_end ? o.row_end : H;
    auto t0 = std::chrono::steady_clock::now();
    std::atomic<uint32_t> next_row(rb);
    std::vector<Counters> cnts(std::max(1, o.threads));
    auto worker = [&](int tid) {
        Counters& cnt = cnts[tid];
        for (;;) {
            uint32_t jj = next_row.fetch_add(1);
            if (jj >= re) break;
            // image rows are stored top-down; the reference's j counts from the bottom
            int j = (int)(H - 1 - jj);
            for (uint32_t i = 0; i < W; i += o.pixel_stride) {
                uint32_t pixel = (uint32_t)j * W + i;
                Rng g;
                g.mode = (int)rd->rng_mode;
                if (g.mode == CRT_RNG_MINSTD) g.ms = Minstd((uint32_t)(rd->seed + pixel));
                g.seed = rd->seed; g.pixel = pixel;
                V3<R> pc(0, 0, 0);
                for (int sj = 0; sj < sq; ++sj) {
                    for (int si = 0; si < sq; ++si) {
                        g.sample = (uint32_t)(sj * sq + si);
                        g.bounce(0);
                        // render kernel sample loop (core/src/gpu.cu:776-790)
                        float r = g.uniform();
                        R u = (R(i) + (R(si) + R(r)) / R(sq)) / R(W);
                        R v = (R(j) + (R(sj) + R(r)) / R(sq)) / R(H);
                        Ray<R> ray = camera_ray(s, u, v, g);
                        V3<R> L = ray_color(s, ray, (int)rd->max_depth, g, cnt);
                        pc = pc + L;
                    }
                }
                float* a = accum + ((size_t)jj * W + i) * 4;
                a[0] = (float)pc.x; a[1] = (float)pc.y; a[2] = (float)pc.z; a[3] = (float)(sq * sq);
            }
        }
    };
    std::vector<std::thread> th;
    int nt = std::max(1, o.threads);
    for (int t = 1; t < nt; ++t) th.emplace_back(worker, t);
    worker(0);
    for (auto& t : th) t.join();
    secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (auto& c : cnts) {
        total.nodes += c.nodes; total.prims += c.prims; total.rays += c.rays; total.retraces += c.retraces;
        total.light_tests += c.light_tests;
    }
    return CRT_OK;
}

}  // namespace orc

extern "C" {

typedef struct oracle_opts {
    int32_t precision;      /* 32: float (refcpu with MINSTD), 64: double (oracle64) */
    int32_t accel;          /* 0: reference-style median BVH, 1: brute force */
    int32_t threads;
    uint32_t row_begin, row_end;  /* rows [begin, end) of the top-down image; 0,0 = all */
    uint32_t pixel_stride;
    int32_t ref_style;      /* 1: refcpu semantics (MT triangles, diagonal quad boxes); 0: product semantics */
} oracle_opts;

typedef struct oracle_stats {
    uint64_t paths, rays, retraces, light_pdf_tests, node_visits, prim_tests;
    double seconds;
} oracle_stats;

int oracle_render(const crt_scene_desc* d, const crt_render_desc* rd, const oracle_opts* opt, float* accum,
                  oracle_stats* st) {
    if (!d || !rd || !opt || !accum) return CRT_ERR_INVALID_ARGUMENT;
    orc::RenderOpts o;
    o.precision = opt->precision; o.accel = opt->accel; o.threads = opt->threads > 0 ? opt->threads : 1;
    o.row_begin = opt->row_begin; o.row_end = opt->row_end; o.pixel_stride = opt->pixel_stride ? opt->pixel_stride : 1;
    o.ref_style = opt->ref_style != 0;
    orc::Counters c; double secs = 0;
    int rc = o.precision == 64 ? orc::render_impl<double>(d, rd, o, accum, c, secs)
                               : orc::render_impl<float>(d, rd, o, accum, c, secs);
    if (st) {
        uint32_t re = o.row_end ? o.row_end : rd->height;
        int sq = (int)std::sqrt((double)rd->spp);
        uint64_t px_per_row = (rd->width + o.pixel_stride - 1) / o.pixel_stride;
        st->paths = (uint64_t)(re - o.row_begin) * px_per_row * (uint64_t)(sq * sq);
        st->rays = c.rays; st->retraces = c.retraces; st->light_pdf_tests = c.light_tests;
        st->node_visits = c.nodes; st->prim_tests = c.prims; st->seconds = secs;
    }
    return rc;
}

/* Closest hit of arbitrary rays: rays n x 8 doubles (o, tmin, d, tmax). */
int oracle_trace(const crt_scene_desc* d, const double* rays, uint64_t n, int32_t camera_rays, int32_t precision,
                 int32_t accel, int32_t ref_style, int32_t* prim, int32_t* inst, double* t, double* uv,
                 double* normal) {
    auto run = [&](auto tag) -> int {
        using R = decltype(tag);
        orc::Scene<R> s;
        int rc = orc::build_scene(s, d, ref_style != 0, accel);
        if (rc) return rc;
        if (accel == 0 && !s.objs.empty()) {
            std::vector<int> ids(s.objs.size());
            for (size_t i = 0; i < ids.size(); ++i) ids[i] = (int)i;
            orc::Minstd brng(1984);
            s.root = orc::build_ref_bvh(s, ids, 0, (int)ids.size(), brng);
        }
        orc::Counters cnt;
        for (uint64_t k = 0; k < n; ++k) {
            const double* q = rays + 8 * k;
            orc::Ray<R> r{{R(q[0]), R(q[1]), R(q[2])}, {R(q[4]), R(q[5]), R(q[6])}, 0};
            orc::Hit<R> h;
            if (orc::world_hit(s, r, R(q[3]), R(q[7]), camera_rays != 0, h, cnt)) {
                const auto& o = s.objs[h.obj];
                prim[k] = o.gid; inst[k] = o.inst; t[k] = (double)h.t;
                if (o.type == orc::OBJ_TRI) { uv[2 * k] = (double)h.bu; uv[2 * k + 1] = (double)h.bv; }
                else { uv[2 * k] = (double)h.u; uv[2 * k + 1] = (double)h.v; }
                if (normal) for (int a = 0; a < 3; ++a) normal[3 * k + a] = (double)h.normal[a];
            } else {
                prim[k] = -1; inst[k] = -1; t[k] = 0; uv[2 * k] = uv[2 * k + 1] = 0;
                if (normal) normal[3 * k] = normal[3 * k + 1] = normal[3 * k + 2] = 0;
            }
        }
        return CRT_OK;
    };
    return precision == 64 ? run(double()) : run(float());
}

/* Pixel-centre camera rays (sample jitter 0.5, no lens) in double: n = W*H rows of 8 doubles. */
int oracle_primary_rays(const crt_scene_desc* d, uint32_t W, uint32_t H, double* rays) {
    orc::Scene<double> s;
    orc::setup_camera(s, d->camera, W);
    for (uint32_t jj = 0; jj < H; ++jj) {
        uint32_t j = H - 1 - jj;
        for (uint32_t i = 0; i < W; ++i) {
            double u = ((double)i + 0.5) / W, v = ((double)j + 0.5) / H;
            orc::V3<double> p = s.llc + u * s.horizontal + v * s.vertical;
            orc::V3<double> o = s.ortho ? p : s.origin;
            orc::V3<double> dir = s.ortho ? -1.0 * s.cw : p - s.origin;
            double* q = rays + 8 * ((size_t)jj * W + i);
            q[0] = o.x; q[1] = o.y; q[2] = o.z; q[3] = 0.00001;
            q[4] = dir.x; q[5] = dir.y; q[6] = dir.z; q[7] = std::numeric_limits<float>::max();
        }
    }
    return CRT_OK;
}

/* Tonemap of the render kernel epilogue (core/src/gpu.cu:792-799, color.cuh:289-309). */
int oracle_tonemap(const float* accum, uint64_t npix, uint32_t spp, uint8_t* rgb) {
    for (uint64_t p = 0; p < npix; ++p) {
        for (int c = 0; c < 3; ++c) {
            float x = accum[4 * p + c];
            x = (x != x) ? 0.0f : x;
            if (spp > 0) { float scale = 1.0f / (float)spp; x *= scale; }
            x = std::sqrt(x);
            float cl = x < 0.0f ? 0.0f : (x > 0.999f ? 0.999f : x);
            rgb[3 * p + c] = (uint8_t)(int)(255.99f * cl);
        }
    }
    return CRT_OK;
}

int oracle_minstd_stream(uint32_t seed, uint32_t n, uint32_t* out, float* u_out) {
    orc::Minstd a(seed), b(seed);
    for (uint32_t i = 0; i < n; ++i) { out[i] = a.next(); if (u_out) u_out[i] = b.uniform_f(); }
    return CRT_OK;
}

int oracle_minstd_uniform_int(uint32_t seed, uint32_t n, int32_t lo, int32_t hi, int32_t* out) {
    orc::Minstd a(seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = a.uniform_int(lo, hi);
    return CRT_OK;
}

int oracle_pcg32_stream(uint64_t initstate, uint64_t initseq, uint32_t n, uint32_t* out) {
    orc::Pcg32 g; g.seed(initstate, initseq);
    for (uint32_t i = 0; i < n; ++i) out[i] = g.next();
    return CRT_OK;
}

int oracle_keyed_stream(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t n, uint32_t* out) {
    orc::Pcg32 g; orc::keyed_seed(g, seed, pixel, sample, bounce);
    for (uint32_t i = 0; i < n; ++i) out[i] = g.next();
    return CRT_OK;
}

/* Reference BVH statistics (depth, nodes) for tests. */
int oracle_ref_bvh_info(const crt_scene_desc* d, uint64_t* nodes, uint32_t* depth) {
    orc::Scene<float> s;
    int rc = orc::build_scene(s, d, true, 0);
    if (rc) return rc;
    std::vector<int> ids(s.objs.size());
    for (size_t i = 0; i < ids.size(); ++i) ids[i] = (int)i;
    orc::Minstd brng(1984);
    if (!ids.empty()) s.root = orc::build_ref_bvh(s, ids, 0, (int)ids.size(), brng);
    *nodes = s.nodes.size();
    // depth of the reference median tree
    std::function<uint32_t(int)> dep = [&](int n) -> uint32_t {
        const auto& nd = s.nodes[n];
        uint32_t a = nd.leaf_l ? 1 : 1 + dep(nd.left);
        uint32_t b = nd.leaf_r ? 1 : 1 + dep(nd.right);
        return a > b ? a : b;
    };
    *depth = s.root >= 0 ? dep(s.root) : 0;
    return CRT_OK;
}

/* ---- entry points of single oracle functions, for the known-answer and closed-form tests ---- */

/* Camera basis of setup_camera (FP64): origin, lower-left corner, horizontal, vertical, u, v, w
 * (3 doubles each) and the lens radius: 22 doubles. */
int oracle_camera_basis(const crt_scene_desc* d, uint32_t width, double* out) {
    orc::Scene<double> s;
    orc::setup_camera(s, d->camera, width);
    const orc::V3<double> v[7] = {s.origin, s.llc, s.horizontal, s.vertical, s.cu, s.cv, s.cw};
    for (int k = 0; k < 7; ++k) { out[3 * k] = v[k].x; out[3 * k + 1] = v[k].y; out[3 * k + 2] = v[k].z; }
    out[21] = s.lens_radius;
    return CRT_OK;
}

/* camera_ray at image-plane points uv[2k], uv[2k+1] (FP64), draws from the keyed PCG stream
 * (seed, pixel k, sample 0, bounce 0): rows of 8 doubles (origin, time, direction, 0). */
int oracle_camera_rays(const crt_scene_desc* d, uint32_t width, uint64_t n, const double* uv, uint64_t seed, double* out) {
    orc::Scene<double> s;
    orc::setup_camera(s, d->camera, width);
    for (uint64_t k = 0; k < n; ++k) {
        orc::Rng g;
        g.mode = CRT_RNG_PCG; g.seed = seed; g.pixel = (uint32_t)k; g.sample = 0;
        g.bounce(0);
        orc::Ray<double> r = orc::camera_ray(s, uv[2 * k], uv[2 * k + 1], g);
        double* q = out + 8 * k;
        q[0] = r.o.x; q[1] = r.o.y; q[2] = r.o.z; q[3] = r.time; q[4] = r.d.x; q[5] = r.d.y; q[6] = r.d.z; q[7] = 0;
    }
    return CRT_OK;
}

/* metal_dir (FP64) for n incoming directions / normals, draws from keyed PCG (seed, k, 0, 0). */
int oracle_metal(uint64_t n, const double* dirs, const double* nrms, double fuzz, uint64_t seed, double* out) {
    for (uint64_t k = 0; k < n; ++k) {
        orc::Rng g;
        g.mode = CRT_RNG_PCG; g.seed = seed; g.pixel = (uint32_t)k; g.sample = 0;
        g.bounce(0);
        orc::V3<double> dd(dirs[3 * k], dirs[3 * k + 1], dirs[3 * k + 2]), nn(nrms[3 * k], nrms[3 * k + 1], nrms[3 * k + 2]);
        orc::V3<double> r = orc::metal_dir(dd, nn, fuzz, g);
        out[3 * k] = r.x; out[3 * k + 1] = r.y; out[3 * k + 2] = r.z;
    }
    return CRT_OK;
}

/* dielectric_dir (FP64) with the draw given per call (draws[k]); reflected[k] = 1 when the ray
 * reflects, drawn[k] = 1 when the draw was consumed. */
int oracle_dielectric(uint64_t n, const double* dirs, const double* nrms, const int32_t* front, double ir,
                      const double* draws, double* out, int32_t* reflected, int32_t* drawn) {
    for (uint64_t k = 0; k < n; ++k) {
        orc::V3<double> dd(dirs[3 * k], dirs[3 * k + 1], dirs[3 * k + 2]), nn(nrms[3 * k], nrms[3 * k + 1], nrms[3 * k + 2]);
        bool refl = false;
        int used = 0;
        orc::V3<double> r = orc::dielectric_dir(dd, nn, front[k] != 0, ir, [&]() { used = 1; return draws[k]; }, &refl);
        out[3 * k] = r.x; out[3 * k + 1] = r.y; out[3 * k + 2] = r.z;
        reflected[k] = refl ? 1 : 0;
        drawn[k] = used;
    }
    return CRT_OK;
}

/* get_sphere_uv at unit-sphere points p (precision 32 or 64). */
int oracle_sphere_uv(uint64_t n, const double* p, const crt_uvmap* map, int32_t precision, double* uv) {
    for (uint64_t k = 0; k < n; ++k) {
        if (precision == 64) {
            double u, v;
            orc::sphere_uv(orc::V3<double>(p[3 * k], p[3 * k + 1], p[3 * k + 2]), *map, u, v);
            uv[2 * k] = u; uv[2 * k + 1] = v;
        } else {
            float u, v;
            orc::sphere_uv(orc::V3<float>((float)p[3 * k], (float)p[3 * k + 1], (float)p[3 * k + 2]), *map, u, v);
            uv[2 * k] = u; uv[2 * k + 1] = v;
        }
    }
    return CRT_OK;
}

/* tex_value of texture `tex` of a scene at (u, v, p): q rows of 5 doubles, rgb rows of 3. */
int oracle_tex_value(const crt_scene_desc* d, int32_t tex, uint64_t n, const double* q, double* rgb) {
    orc::Scene<double> s;
    int rc = orc::build_scene(s, d, false, 1);
    if (rc) return rc;
    if (tex < 0 || tex >= (int32_t)s.texs.size()) return CRT_ERR_INVALID_ARGUMENT;
    for (uint64_t k = 0; k < n; ++k) {
        const double* a = q + 5 * k;
        orc::V3<double> c = orc::tex_value(s, tex, a[0], a[1], orc::V3<double>(a[2], a[3], a[4]));
        rgb[3 * k] = c.x; rgb[3 * k + 1] = c.y; rgb[3 * k + 2] = c.z;
    }
    return CRT_OK;
}

/* Light PDF value at point o towards direction v (for Monte-Carlo normalisation tests). */
int oracle_light_pdf(const crt_scene_desc* d, const double* o, const double* v, uint64_t n, double* out) {
    orc::Scene<double> s;
    int rc = orc::build_scene(s, d, false, 1);
    if (rc) return rc;
    orc::Counters c;
    for (uint64_t k = 0; k < n; ++k) {
        orc::V3<double> oo(o[3 * k], o[3 * k + 1], o[3 * k + 2]), vv(v[3 * k], v[3 * k + 1], v[3 * k + 2]);
        out[k] = orc::lights_pdf_value(s, oo, vv, c);
    }
    return CRT_OK;
}

/* Light sample directions (lights_random, FP64) from points o, draws from the keyed PCG stream
 * (seed, pixel k, sample 0, bounce 0). */
int oracle_light_sample(const crt_scene_desc* d, const double* o, uint64_t n, uint64_t seed, double* out) {
    orc::Scene<double> s;
    int rc = orc::build_scene(s, d, false, 1);
    if (rc) return rc;
    for (uint64_t k = 0; k < n; ++k) {
        orc::Rng g;
        g.mode = CRT_RNG_PCG; g.seed = seed; g.pixel = (uint32_t)k; g.sample = 0;
        g.bounce(0);
        orc::V3<double> r = orc::lights_random(s, orc::V3<double>(o[3 * k], o[3 * k + 1], o[3 * k + 2]), g);
        out[3 * k] = r.x; out[3 * k + 1] = r.y; out[3 * k + 2] = r.z;
    }
    return CRT_OK;
}

}  // extern "C"
