// scene_build.h - host-side flattening of a crt_scene_desc into the HBM layout of crt_layout.h,
// including the binned-SAH BVH2 build (two levels when the scene has instances).
//
// Replaces the reference's device-side scene construction: load_scene<<<1,1>>>
// (core/src/gpu.cu:481-751, object graph + name lookups at :132-474) and the single-thread BVH
// (core/src/misc/bvh_node.cuh:63-108: random axis, O(n^2) sort, count-median split).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "cortex_rt.h"

namespace crt {

struct F4 { float x, y, z, w; };
struct F2 { float x, y; };
struct U2 { uint32_t x, y; };
struct U4 { uint32_t x, y, z, w; };

struct BuildStats {
    uint64_t nodes = 0, leaves = 0, prims = 0;
    uint32_t depth = 0;        // max interior depth of the top level (+ deepest BLAS when instanced)
    double build_ms = 0.0;
    double sah_cost = 0.0;     // SAH cost of the top-level tree (C_trav = 1, C_isect = 1)
};

struct HostScene {
    std::vector<F4> nodes;        // 4 per interior node
    std::vector<F4> tris;         // 3 per triangle
    std::vector<F4> tri_normals;  // 3 per triangle, only when a mesh has normals
    std::vector<F2> tri_uvs;      // 3 per triangle, only when a mesh has uvs
    std::vector<F4> spheres;      // 2 per sphere: (centre, radius), (global id, material, -, -) as bits
    std::vector<U4> sphere_aux;   // host only: (global id, material) per sphere in leaf order
    std::vector<F2> sphere_map;   // 3 per sphere: scale, offset, repeat
    std::vector<F4> quads;        // 5 per quad / light
    std::vector<F4> insts;        // 4 per instance
    std::vector<F4> mats;         // 2 per material (scene materials, then one per area light), then the
                                  // shapes' uv mappings (2 per shape from shape_map_base)
    uint32_t shape_map_base = 0;  // float4 index of the first shape's uv mapping in mats
    std::vector<F4> texs;         // 2 per texture
    std::vector<uint8_t> texels;
    std::vector<uint32_t> light_refs;   // per light in world order: quad slot, or CRT_LIGHT_SPHERE | sphere slot
    uint32_t root = 0;
    uint32_t bvh_width = 4;       // 2: 64-B binary nodes, 4: 128-B four-wide SoA nodes
    uint32_t leaf_size = 1;       // max primitives per leaf the tree was built with
    float background[3] = {0, 0, 0};
    crt_camera camera{};
    uint32_t num_lights = 0;        // omni + area + spot lights
    uint32_t num_spheres = 0;       // sphere records (sphere lights and spheres)
    uint32_t num_quads = 0;         // the scene's quads (not its area lights)
    uint32_t num_quad_records = 0;  // quad records (area lights and quads)
    uint32_t num_instances = 0;
    uint32_t num_shapes = 0;        // analytic shapes (records in the quad array)
    uint64_t num_world_tris = 0;
    uint64_t num_node_records = 0;  // BVH nodes in the device allocation (BLAS nodes included)
    uint64_t num_tri_records = 0;   // triangle records in the device allocation (BLAS triangles included)
    bool device_built = false;    // BVH and records built on the device (no host copies of them)
    BuildStats stats;
    uint64_t device_bytes() const;
    // Nodes, triangle, instance, sphere and quad records share one device allocation (in that
    // order, each at a 256-byte aligned offset); the traversal addresses them with 32-bit byte
    // offsets from the nodes.
    static uint64_t align256(uint64_t b) { return (b + 255u) & ~(uint64_t)255u; }
    uint64_t tri_offset() const { return align256(nodes.size() * sizeof(F4)); }
    uint64_t inst_offset() const { return align256(tri_offset() + tris.size() * sizeof(F4)); }
    uint64_t sphere_offset() const { return align256(inst_offset() + insts.size() * sizeof(F4)); }
    uint64_t quad_offset() const { return align256(sphere_offset() + spheres.size() * sizeof(F4)); }
    uint64_t geometry_bytes() const { return quad_offset() + quads.size() * sizeof(F4); }
};

// Largest geometry allocation of one scene (32-bit byte offsets in the traversal).
constexpr uint64_t kMaxGeometryBytes = (1ull << 32) - 256u;

// Checks the descriptor and fills the texture and material tables, camera, background, BVH width and
// leaf size of S (the part of the flattening shared by the host and the device builds).
int prepare_scene_tables(const crt_scene_desc* d, uint32_t leaf_size, uint32_t bvh_width, HostScene& S, std::string& err);

// Primitive records in source order (area lights then quads; sphere lights then spheres; mesh
// triangles) for the device BVH build of a one-level scene; the same record layouts as HostScene,
// before they are put in leaf order.
struct SourceRecords {
    std::vector<F4> quads;        // 5 per quad / area light / shape, lights first
    std::vector<F4> quad_boxes;   // 2 per quad record: (lo, -), (hi, -)
    std::vector<uint32_t> quad_gid;
    uint32_t num_lights = 0;      // omni + area + spot lights
    std::vector<uint32_t> light_src;   // world light order: quad source index, or CRT_LIGHT_SPHERE | sphere source index
    std::vector<F4> spheres;      // 2 per sphere (sphere lights first)
    std::vector<F2> sphere_map;   // 3 per sphere
    std::vector<U4> sphere_aux;   // (global id, material) per sphere
    std::vector<F4> tris;         // 3 per triangle
    std::vector<F4> tri_normals;  // 3 per triangle when a mesh has normals
    std::vector<F2> tri_uvs;      // 3 per triangle when a mesh has uvs
    uint32_t tri_gid0 = 0;        // global id of the first triangle (consecutive after it)
};
int flatten_sources(const crt_scene_desc* d, uint32_t shape_map_base, int threads, SourceRecords& out, std::string& err);

// Flattens and builds. leaf_size = max primitives per leaf (instances always get their own leaf);
// bvh_width = 2 or 4 children per node (the binary SAH tree is collapsed for 4).
int build_host_scene(const crt_scene_desc* d, uint32_t leaf_size, uint32_t bvh_width, int threads, HostScene& out,
                     std::string& err);

// Image texture file (PNG or JPEG; image_io.cpp): rows top-down, C = 3 (RGB8) or 4 (RGBA8).
bool read_image(const std::string& path, std::vector<uint8_t>& out, uint32_t& W, uint32_t& H, uint32_t& C, std::string& err);

// Structural BVH invariants (containment, exact leaf coverage, reachability).
int check_host_scene(const HostScene& S, std::string& err);

}  // namespace crt
