// crt_layout.h - HBM layout of scenes, requests and wavefront path state, shared by the host
// builder (C++) and the gfx950 kernels (HIP). Everything here is POD and 16-byte aligned where it
// is read as float4.
//
// Scene (one per crt_scene_id, resident until released):
//   nodes   : BVH2, 4 x float4 (64 B) per interior node; both children's boxes live in the parent
//             n0 = (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y)
//             n1 = (c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y)
//             n2 = (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z)
//             n3 = (child0, child1, -, -) as int: >= 0 interior node, < 0 inline leaf reference
//             (CRT_LEAF_REF below: type, count and first record of the leaf's primitives)
//           or BVH4, 8 x float4 (128 B) per interior node, one float4 per axis bound over the four
//             children: lo.x hi.x lo.y hi.y lo.z hi.z, then (child0..child3) as int, then padding
//   tris    : 3 x float4 per triangle in leaf order: (v0, gid) (v1, material) (v2, flags)
//   spheres : 2 x float4 per sphere in leaf order: (centre, radius) (gid, material, -, -) as bits
//   quads   : 5 x float4 per quad/area light: (Q, D) (u, flags) (v, area) (n, gid) (w, material)
//             or per analytic shape (disk, cylinder, cone; flags CRT_QF_DISK.. below), whose hit runs
//             in its own frame: local = A x + b for a world point x, with
//             (A row 0, b.x) (radius, height, -, flags) (A row 1, b.y) (A row 2, gid)
//             (b.z, uv-mapping slot in the material table as uint, -, material)
//   insts   : 4 x float4 per instance: three rows of world->object [R|t] and
//             (blas_root, instance id, -, -) as int
// Materials (2 x float4), textures (2 x float4) and one texel byte pool complete the scene.
#pragma once
#include <stdint.h>

#define CRT_PRIM_TRI 0u
#define CRT_PRIM_SPHERE 1u
#define CRT_PRIM_QUAD 2u
#define CRT_PRIM_INST 3u

// Child references in interior nodes: >= 0 an interior node; < 0 a leaf, written inline so the
// traversal reads the leaf's primitive records without an indirection:
//   bit 31 | type << 29 | (count - 1) << 26 | first   (first + count <= 2^26, 1 <= count <= 8)
#define CRT_LEAF_MAX_COUNT 8u
#define CRT_LEAF_MAX_FIRST (1u << 26)
#define CRT_LEAF_REF(type, first, count) ((int32_t)(0x80000000u | ((type) << 29) | (((count)-1u) << 26) | (first)))
#define CRT_LEAF_REF_TYPE(r) (((uint32_t)(r) >> 29) & 3u)
#define CRT_LEAF_REF_COUNT(r) ((((uint32_t)(r) >> 26) & 7u) + 1u)
#define CRT_LEAF_REF_FIRST(r) ((uint32_t)(r)&0x03ffffffu)

// quad flags (float4 #1 .w as uint)
#define CRT_QF_LIGHT 1u        // centred parametrisation, interior in [-0.5, 0.5]^2, uv shifted by 0.5
#define CRT_QF_INVISIBLE 2u    // hidden from camera rays (directional_light.cuh:106-109)
#define CRT_QF_ONE_SIDED 4u    // back faces trigger the pass-through re-trace
#define CRT_QF_DISK 8u         // analytic shapes (records in the quad array)
#define CRT_QF_CYLINDER 16u
#define CRT_QF_CONE 32u
#define CRT_QF_SHAPE (CRT_QF_DISK | CRT_QF_CYLINDER | CRT_QF_CONE)
#define CRT_QF_UNIT_NORMAL 64u // shape under a scale: the world normal is renormalised (scale.cuh:50-56)
// triangle flags (float4 #2 .w as uint)
#define CRT_TF_SMOOTH 1u
#define CRT_TF_UV 2u
// sphere flags (float4 #1 .z as uint): sphere lights (omni_light.cuh, spot_light.cuh)
#define CRT_SF_INVISIBLE 1u    // hidden from camera rays (omni_light.cuh:62-65)
#define CRT_SF_SPOT 2u         // spot light: back faces trigger the pass-through re-trace
// light table entry: leaf-ordered quad index, or CRT_LIGHT_SPHERE | leaf-ordered sphere index
#define CRT_LIGHT_SPHERE 0x80000000u

// hit.prim encoding: type << 30 | leaf-ordered index; CRT_MISS = no hit
#define CRT_MISS 0xffffffffu
#define CRT_HIT_TYPE(p) ((p) >> 30)
#define CRT_HIT_INDEX(p) ((p)&0x3fffffffu)
// hit.inst: instance record index (leaf order) or CRT_NO_INST; bit 31 marks a pass-through re-trace
// result, bit 30 one behind a visible spot light's back face (which adds white, diffuse_spot_light.cuh:41-44)
#define CRT_NO_INST 0x3fffffffu
#define CRT_PASSTHROUGH 0x80000000u
#define CRT_PASSTHROUGH_WHITE 0x40000000u
#define CRT_INST_MASK 0x3fffffffu

// material record (2 x float4)
//  m0 = (albedo.rgb, param)   param = fuzz (metal) | ior (dielectric) | intensity (light)
//  m1 = (type, texture, flags, -) as int
// texture record (2 x float4)
//  t0 = (color.rgb | -, scale)
//  t1 = (type | width<<? ...) -> as ints: (type, a, b, c) with
//       SOLID: -, CHECKER: (type, even, odd, -), IMAGE: (type, width, height, pool offset / 4 dwords)
struct crt_dev_scene {
    const void* nodes;      // float4
    const void* tris;       // float4; tris, insts, spheres and quads follow the nodes in one allocation
                            // (at most 4 GiB together)
    const void* tri_normals;// float4 x3 per triangle (optional)
    const void* tri_uvs;    // float2 x3 per triangle (optional)
    const void* spheres;    // float4 x2 per sphere
    const void* sphere_map; // float2 x3 per sphere (scale, offset, repeat)
    const void* quads;      // float4
    const void* insts;      // float4
    const void* mats;       // float4
    const void* texs;       // float4
    const uint8_t* texels;  // byte pool (rows top-down, RGB(A)8)
    const uint32_t* lights;       // light table in world-list order (CRT_LIGHT_SPHERE entries for sphere lights)
    uint32_t num_lights;
    uint32_t root;          // root node index
    uint32_t num_quads;
    uint32_t bvh_width;
    float background[3];
    float pad1;
};

// Per-request parameters read by the generate / shade / resolve kernels.
struct crt_dev_request {
    // camera (core/src/cameras/perspective_camera.cuh:41-113, orthographic_camera.cuh:57-142),
    // computed on the host in FP32; ortho_dir = -w for the orthographic camera's parallel rays
    float origin[3], llc[3], horizontal[3], vertical[3], cu[3], cv[3], ortho_dir[3];
    float lens_radius, time0, time1;
    uint32_t width, height, sqrt_spp, spp, max_depth;
    uint32_t rng_mode;
    uint32_t scene;          // index into the scene table
    uint64_t seed;
    // pass work: pixel range x sample range, laid out pixel-major from path_base (a pixel's
    // samples consecutive)
    uint32_t path_base;      // first path slot of this request in the pass
    uint32_t pix_begin, pix_count;
    uint32_t sample_begin, sample_count;
    float* accum;            // float4 per pixel (rgb sums, sample count), rows top-down
    uint32_t* rng_state;     // MINSTD: per-pixel engine state carried across samples
    uint32_t ortho;          // 1: orthographic camera
    uint32_t pad;
};

// Traversal statistics (counting kernel only)
struct crt_dev_counters {
    unsigned long long node_visits, leaf_visits, tri_tests, sphere_tests, quad_tests, inst_visits;
    unsigned long long rays, retraces, light_tests, stack_overflow;
    unsigned long long exact;   // queries re-traced with the double-precision triangle test
};
