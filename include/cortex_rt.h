/*
 * cortex_rt.h - C-ABI of the MI355X (gfx950) render hot path.
 *
 * This is the drop-in boundary for the reference renderer's one render entry point
 *
 *   void launchGPU(const sceneConfig&, int width, int height, int spp, int max_depth,
 *                  int tx, int ty, const char* filepath, bool quietMode);
 *
 * declared at core/src/renderers/gpu_cuda_renderer.h:17 and defined at core/src/gpu.cu:2127-2149.
 * The reference takes a C++ object graph (sceneConfig, core/src/scenes/scene_config.h:452-460) whose
 * materials and textures are linked by name and rebuilt on the device by one thread
 * (load_scene<<<1,1>>>, core/src/gpu.cu:481-751). Here the caller hands over a flat, POD
 * description (indices instead of names, uint32 counts instead of uint8 ones); the library copies
 * it, builds the BVH on the host and keeps everything resident in HBM until the scene is released.
 *
 * Conventions
 *  - Every function returns a crt_status; 0 is success. crt_last_error() gives the message of the
 *    last failure on the calling thread.
 *  - All pointers in the descriptors are host pointers; the library copies what it needs before
 *    returning, so the caller may free its buffers as soon as the call returns.
 *  - Images are row-major with row 0 at the TOP (the reference flips rows when it writes the PNG,
 *    core/src/gpu.cu:2098-2107; the float accumulation buffer uses the same top-down order).
 */
#ifndef CORTEX_RT_H
#define CORTEX_RT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CRT_ABI_VERSION 3u

/* ------------------------------------------------------------------ status codes */
typedef enum crt_status {
    CRT_OK = 0,
    CRT_ERR_NO_DEVICE = 1,        /* reference: "No GPU found" print + return (core/src/gpu.cu:2131-2134) */
    CRT_ERR_OUT_OF_MEMORY = 2,    /* hipMalloc failed */
    CRT_ERR_OVER_BUDGET = 3,      /* the request does not fit the context's HBM budget */
    CRT_ERR_INVALID_ARGUMENT = 4,
    CRT_ERR_INVALID_SCENE = 5,    /* dangling material/texture/mesh index, NaN geometry, ... */
    CRT_ERR_IO = 6,
    CRT_ERR_CANCELLED = 7,
    CRT_ERR_NOT_READY = 8,        /* crt_poll / crt_try_result: request still running */
    CRT_ERR_HIP = 9,              /* any other HIP runtime error (reference: exit(99), core/src/gpu.cu:119-129) */
    CRT_ERR_INTERNAL = 10
} crt_status;

/* ------------------------------------------------------------------ scene description */

/* Perspective camera, core/src/cameras/perspective_camera.cuh:41-119 (initialize + get_ray), or
 * orthographic camera, core/src/cameras/orthographic_camera.cuh:57-142 (cameraCfg.isOrthographic,
 * core/src/gpu.cu:696-705). The aspect ratio is the camera's own (reference: -ratio CLI flag), not
 * width/height. */
typedef struct crt_camera {
    float lookfrom[3];
    float lookat[3];
    float vup[3];
    float vfov_deg;      /* perspective only */
    float aspect_ratio;
    float aperture;      /* lens radius = aperture / 2 (the orthographic camera offsets its ray origins by
                            the lens sample too) */
    float focus_dist;    /* perspective: primary ray directions have length ~focus_dist (not normalized) */
    float time0, time1;  /* ray time = time0 + U*(time1-time0); the reference passes 0 and 1 */
    uint32_t orthographic;  /* 1: parallel rays along lookat - lookfrom (unit length) from a viewport of
                               height 2 * ortho_height and width 2 * ortho_height * W / int(W / aspect) at
                               lookfrom, W = the render width */
    float ortho_height;
} crt_camera;

/* Material type codes follow MaterialTypeID (core/src/materials/material.cuh:8-19). */
enum {
    CRT_MAT_LAMBERTIAN = 1,      /* core/src/materials/lambertian.cuh:63-143 */
    CRT_MAT_METAL = 2,           /* core/src/materials/metal.cuh:52-65 */
    CRT_MAT_DIELECTRIC = 3,      /* core/src/materials/dielectric.cuh:32-63 */
    CRT_MAT_ISOTROPIC = 4,       /* core/src/materials/isotropic.cuh: albedo from its texture, directions from
                                    the mixture of the lights and the uniform sphere (sphere_pdf), scattering
                                    pdf 1 / (4 pi) */
    CRT_MAT_OREN_NAYAR = 5,      /* core/src/materials/oren_nayar.cuh: fuzz = roughness, ior = albedo_temp;
                                    attenuation light0.rgb * light0.intensity * albedo * albedo_temp / pi *
                                    (A + B alpha beta) with the outgoing cosine 0 (its skip_pdf_ray is unset),
                                    directions from the lights / uniform-sphere mixture, scattering pdf cos / pi */
    CRT_MAT_DIFFUSE_LIGHT = 8,   /* core/src/materials/diffuse_light.cuh:45-67 */
    CRT_MAT_SPOT_LIGHT = 9       /* diffuse_spot_light, core/src/materials/diffuse_spot_light.cuh:33-61: the
                                    material of a spot light (crt_sphere_light), not for crt_material */
};
/* crt_material.flags */
#define CRT_MATF_ONE_SIDED 1u    /* diffuse_light "directional": back faces emit colour(0,0,0,0) */

typedef struct crt_material {
    uint32_t type;
    int32_t texture;     /* index into textures[], or -1 for the constant albedo below */
    float albedo[3];
    float fuzz;          /* metal; clamped to <= 1 at upload like the reference constructor */
    float ior;           /* dielectric index of refraction */
    float intensity;     /* diffuse light: emitted = albedo * intensity */
    uint32_t flags;
} crt_material;

enum {
    CRT_TEX_SOLID = 0,   /* core/src/textures/solid_color_texture.cuh:34-37 */
    CRT_TEX_CHECKER = 1, /* 3-D spatial checker, core/src/textures/checker_texture.cuh:33-42 */
    CRT_TEX_IMAGE = 2    /* nearest texel, V flip, clamp, core/src/textures/image_texture.cuh:54-71 */
};

typedef struct crt_texture {
    uint32_t type;
    float color[3];          /* SOLID */
    float scale;             /* CHECKER: cell size (the reference stores 1/scale) */
    int32_t even, odd;       /* CHECKER: texture indices (must not be checkers themselves) */
    uint32_t width, height;  /* IMAGE */
    uint32_t channels;       /* IMAGE: 3 (RGB8) or 4 (RGBA8); only RGB is sampled */
    const uint8_t* pixels;   /* IMAGE: width*height*channels bytes, row 0 at the top */
} crt_texture;

/* Texture-coordinate mapping of spheres (uvmapping, core/src/utilities/uvmapping.cuh:8-60). */
typedef struct crt_uvmap {
    float scale_u, scale_v, offset_u, offset_v, repeat_u, repeat_v;
} crt_uvmap;

typedef struct crt_sphere {      /* core/src/primitives/sphere.cuh:111-166 */
    float center[3];
    float radius;
    uint32_t material;
    crt_uvmap mapping;
} crt_sphere;

typedef struct crt_quad {        /* core/src/primitives/quad.cuh:56-170; corner + two edges */
    float corner[3];
    float u[3];
    float v[3];
    uint32_t material;
} crt_quad;

/* Quad area light centred at `center` (directional_light, core/src/lights/directional_light.cuh:45-172).
 * Emission = rgb * intensity from the face that u x v points to; invisible lights are skipped by
 * camera rays only (depth == max_depth, directional_light.cuh:106-109). */
typedef struct crt_area_light {
    float center[3];
    float u[3];
    float v[3];
    float rgb[3];
    float intensity;
    uint32_t invisible;
} crt_area_light;

/* Spherical light (core/src/lights/omni_light.cuh:18-93, spot_light.cuh:21-117): a sphere hit like
 * sphere::hit, hidden from camera rays when invisible, sampled by the lights' mixture pdf with
 * random_to_sphere over its cone of directions (pdf 1 / (2 pi (1 - cos theta_max))).
 * OMNI: diffuse_light(rgb, intensity, non-directional): rgb * intensity from both faces.
 * SPOT: diffuse_spot_light: a ray meeting the front face along unit direction D emits
 * rgb * intensity * cos^falloff where cos = D . unit(direction) >= cos(cutoff_deg), and white (1,1,1)
 * where cos is below it; a ray from inside (back face) is passed through like a one-sided light's
 * back face (camera.cuh:179-183), adding white when the light is visible and nothing when invisible. */
enum { CRT_LIGHT_OMNI = 0, CRT_LIGHT_SPOT = 1 };
typedef struct crt_sphere_light {
    uint32_t kind;         /* CRT_LIGHT_OMNI or CRT_LIGHT_SPOT */
    float center[3];
    float radius;
    float rgb[3];
    float intensity;
    uint32_t invisible;
    float direction[3];    /* SPOT: cone axis (any length) */
    float cutoff_deg;      /* SPOT: cone half-angle in degrees */
    float falloff;         /* SPOT: exponent of the cosine */
} crt_sphere_light;

/* Primitive transform (rt::transform, core/src/misc/transform.cuh): translate, rotate (degrees
 * about x, y, z) and scale, each present or not (CRT_XF_* bits in `has`), applied as the reference's
 * translate / rotate / scale wrappers see a primitive (scene_factory.cuh:710-725): the surface is the
 * primitive under p -> S Rz(z) Ry(y) Rx(x) (p + T). A sphere under a non-uniform scale (an ellipsoid,
 * outside the path) keeps its radius and is counted unsupported. */
#define CRT_XF_TRANSLATE 1u
#define CRT_XF_ROTATE 2u
#define CRT_XF_SCALE 4u
typedef struct crt_transform {
    uint32_t has;
    float translate[3];
    float rotate_deg[3];
    float scale[3];
} crt_transform;

/* Analytic shapes (core/src/primitives/disk.cuh:44-82, cylinder.cuh:50-103, cone.cuh:44-128),
 * built as load_scene builds them (core/src/gpu.cu:565-595) and placed by `transform`:
 * DISK     the disk |(x, z) - (cx, cz)| <= radius in the plane y = cy, normal +y (both faces);
 *          `height` only sizes its box
 * CYLINDER the open tube of `radius` around the vertical axis through `center`, cy <= y <= cy + height
 *          (no caps); its normal is (p - center) / radius, not renormalised, as the reference has it
 * CONE     the side of the cone with its base circle of `radius` at y = cy and its apex at
 *          y = cy + height (no base cap)
 * Texture coordinates follow get_disk_uv / get_cylinder_uv / get_cone_uv
 * (core/src/utilities/uvmapping.cuh:285-368) under `mapping`. Shapes are never lights. */
enum { CRT_SHAPE_DISK = 0, CRT_SHAPE_CYLINDER = 1, CRT_SHAPE_CONE = 2 };
typedef struct crt_shape {
    uint32_t kind;
    float center[3];
    float radius, height;
    uint32_t material;
    crt_uvmap mapping;
    crt_transform transform;
} crt_shape;

/* Indexed triangle mesh (Moller-Trumbore in the reference, core/src/primitives/triangle.cuh:130-185;
 * watertight here). normals/uvs are optional per-vertex arrays (NULL = flat normal / zero uv). */
typedef struct crt_mesh {
    const float* positions;        /* 3 floats per vertex */
    uint32_t num_vertices;
    const uint32_t* indices;       /* 3 per triangle */
    uint32_t num_triangles;
    const float* normals;          /* optional, 3 per vertex: smooth shading */
    const float* uvs;              /* optional, 2 per vertex */
    uint32_t material;             /* material of every triangle unless tri_materials is given */
    const uint32_t* tri_materials; /* optional, 1 per triangle */
} crt_mesh;

/* Rigid or affine placement of a mesh (the reference's translate/rotate/scale wrappers,
 * core/src/primitives/{translate,rotate,scale}.cuh, applied per primitive). Row-major 3x4
 * object-to-world matrix [R|t]. When a scene has instances, meshes[] are only reachable through
 * instances (two-level BVH); otherwise every mesh is placed in world space as given. */
typedef struct crt_instance {
    uint32_t mesh;
    float object_to_world[12];
    int32_t material_override;     /* -1: keep the mesh's materials */
} crt_instance;

typedef struct crt_scene_desc {
    uint32_t abi_version;          /* CRT_ABI_VERSION */
    crt_camera camera;
    float background[3];           /* miss colour (imageCfg.background.rgb, core/src/gpu.cu:741) */
    const crt_material* materials; uint32_t num_materials;
    const crt_texture* textures;   uint32_t num_textures;
    const crt_area_light* lights;  uint32_t num_lights;
    const crt_sphere* spheres;     uint32_t num_spheres;
    const crt_quad* quads;         uint32_t num_quads;
    const crt_mesh* meshes;        uint32_t num_meshes;
    const crt_instance* instances; uint32_t num_instances;
    const crt_sphere_light* sphere_lights; uint32_t num_sphere_lights;
    const crt_shape* shapes;       uint32_t num_shapes;
} crt_scene_desc;

/* Global primitive numbering used by first-hit queries (mirrors the reference's world-list order,
 * lights first, then spheres, quads, shapes, mesh triangles: core/src/gpu.cu:504-640):
 *   lights [0, L), spheres [L, L+S), quads [L+S, L+S+Q), shapes [L+S+Q, L+S+Q+H),
 *   triangles [L+S+Q+H, ...) in mesh order.
 * The lights are the OMNI sphere lights, then the area lights, then the SPOT sphere lights (each
 * in array order), as load_scene adds omni, directional and spot lights (core/src/gpu.cu:504-523);
 * the light sampled by the estimator's mixture pdf is the first of them (hittable_list::random's
 * get_int always returns its lower bound, core/src/misc/gpu_randomizer.cuh:29-36).
 * For instanced scenes the triangle id is the mesh-local id and the instance is reported apart. */

/* ------------------------------------------------------------------ rendering */
enum {
    CRT_RNG_PCG = 0,     /* PCG32 stream keyed by (seed, pixel, sample, bounce): order independent */
    CRT_RNG_MINSTD = 1   /* reference stream: one minstd_rand per pixel seeded seed+pixel
                            (core/src/gpu.cu:770, seed 2580 at :2080); samples run in order */
};

typedef struct crt_render_desc {
    uint32_t width, height;
    uint32_t spp;          /* the loop runs floor(sqrt(spp))^2 samples and divides by spp,
                              as the reference does (core/src/gpu.cu:776-792) */
    uint32_t max_depth;
    uint64_t seed;         /* PCG: stream seed; MINSTD: the per-pixel seed base (2580 in the reference) */
    uint32_t rng_mode;     /* CRT_RNG_* */
    uint32_t sample_begin; /* first sample index of this call (progressive rendering) */
    uint32_t sample_count; /* 0 = all remaining samples */
    uint32_t flags;        /* CRT_RENDER_* bits */
} crt_render_desc;

/* crt_render_desc.flags */
enum {
    CRT_RENDER_TIME_STAGES = 1u,  /* time the closest-hit launches with HIP events (crt_stats.extend_ms) */
    CRT_RENDER_MEGAKERNEL = 2u,   /* blocking renders only: one kernel launch renders the request, one thread per
                                     pixel running its samples and bounces in turn (the reference's render kernel
                                     design, core/src/gpu.cu:756-801) instead of the wavefront passes. Same
                                     estimator and streams: the image agrees with the wavefront render. */
    CRT_RENDER_IMAGE_ONLY = 4u    /* served requests: only the RGB8 image is copied back at completion;
                                     crt_wait then returns no accumulation buffer (CRT_ERR_INVALID_ARGUMENT
                                     if one is asked for). Blocking renders skip it when crt_image.accum
                                     is NULL. */
};

typedef struct crt_image {
    uint8_t* rgb8;         /* optional: width*height*3, tonemapped like core/src/gpu.cu:792-799 */
    float* accum;          /* optional: width*height*4 float radiance sums (rgb, sample count) */
} crt_image;

typedef struct crt_stats {
    uint64_t paths;            /* camera paths traced */
    uint64_t rays;             /* closest-hit queries: one per path segment plus every pass-through re-trace
                                  (the estimator has no shadow rays) */
    uint64_t retraces;         /* pass-through re-traces behind one-sided lights (core/src/cameras/camera.cuh:180-183),
                                  included in rays */
    uint64_t light_pdf_tests;  /* analytic light-PDF intersection tests */
    uint64_t node_visits;      /* interior nodes fetched (counting kernels only, else 0) */
    uint64_t prim_tests;       /* primitive intersection tests (counting kernels only, else 0) */
    double gpu_ms;             /* device time of the render passes (HIP events) */
    double extend_ms;          /* device time spent in the traversal kernel */
    double host_ms;            /* wall time of the call */
} crt_stats;

/* ------------------------------------------------------------------ context */
typedef struct crt_context crt_context;
typedef uint32_t crt_scene_id;
typedef uint64_t crt_request_id;

typedef struct crt_config {
    int32_t device;             /* HIP device ordinal */
    uint64_t hbm_budget_bytes;  /* admission budget for scenes + requests; 0 = 90% of free memory */
    uint32_t max_paths;         /* live path slots of the streaming passes (path state pools, 164 B each);
                                   0 = 64M */
    uint32_t block_size;        /* threads per block of the stage kernels; 0 = default (256) */
    uint32_t count_work;        /* 1: use the counting traversal kernel (fills node_visits/prim_tests) */
    uint32_t leaf_size;         /* BVH max primitives per leaf (1-8); 0 = default (1) */
    uint32_t bvh_width;         /* children per BVH node, 2 or 4; 0 = default (4) */
    uint32_t extend_kernel;     /* closest-hit stage (CRT_EXTEND_*): 0 = default, 1 = grid-stride (each lane
                                   traces its rays to the end, the wave waits for its slowest lane),
                                   2 = per-lane refill: a lane whose ray is done takes the next ray of its
                                       wave's contiguous share of the queue (no atomics),
                                   3 = persistent waves taking chunks of 64-ray groups from a work counter */
    uint32_t refill_min;        /* mode 2: idle lanes of a wave before they take new rays (default 16);
                                   mode 3: 64-ray groups per fetch; 0 = default */
    uint32_t pass_paths;        /* path ids per pass (16 B of radiance each); 0 = min(2^27, 16 * max_paths)
                                   for blocking renders, min(2^29, 32 * max_paths) for the serving loop */
    uint32_t grid_per_cu;       /* workgroups per CU of the closest-hit launches (256 threads each); 0 = 24 */
    uint32_t shade_grid_per_cu; /* workgroups per CU of the regenerate / shade launches; 0 = 16 */
    uint32_t sub_passes;        /* concurrent sub-passes (one stream each) of a pass, 1-4; 0 = 2 */
    uint32_t graphs;            /* serving loop iterations replayed from hipGraphs: 0 = default (on), 1 = on,
                                   2 = off (every iteration launched eagerly) */
    uint32_t stack_lds;         /* traversal stack: LDS ring entries per lane behind the register top entry,
                                   8, 16 or 32; 0 = 16 */
    uint32_t ray_sort;          /* 1: the shade stage groups each workgroup's surviving paths by direction
                                   octant (coherent waves in the next closest-hit launch); 0 / 2: off */
    uint32_t bvh_builder;       /* CRT_BVH_*: where crt_scene_upload builds the BVH; 0 = host */
    uint32_t batch_group_paths; /* crt_render_batch: a request with at least this many path ids left gets
                                   passes of its own scene (single-scene kernel, one tree in the caches);
                                   smaller remainders share passes; 0 = 2^22, 0xffffffff = always share */
    uint32_t serve_order;       /* CRT_SERVE_*: the order in which the serving loop feeds admitted requests
                                   into its passes; 0 = CRT_SERVE_SMALL_THEN_OLDEST */
    uint32_t serve_small_paths; /* CRT_SERVE_SMALL_THEN_OLDEST: requests with at most this many path ids left
                                   go first; 0 = 2^22 */
} crt_config;

/* crt_config.serve_order. SMALL_THEN_OLDEST: requests with at most serve_small_paths path ids left
 * first (in admission order), then the others oldest first, each taking as much of a pass as its
 * backlog bound allows: an overloaded loop then drains its large requests in arrival order, which
 * bounds the largest latency, while small requests pass them. ROUND_ROBIN: one segment per request
 * in turn (every request progresses; large requests finish last, whenever they arrived). Images do
 * not depend on the order. */
enum { CRT_SERVE_SMALL_THEN_OLDEST = 1, CRT_SERVE_ROUND_ROBIN = 2 };

/* crt_config.bvh_builder. HOST: binned-SAH tree on the CPU threads, then one upload. DEVICE: linear
 * BVH on the GPU (Morton codes, radix sort, radix tree, bottom-up boxes, collapse to the BVH width)
 * from the uploaded primitives: a build in milliseconds for scenes that arrive with a request, at a
 * higher traversal cost than the SAH tree. One-level scenes of two primitives or more; instanced
 * and smaller scenes are built on the host either way. Closest hits, and so images, do not depend
 * on the tree (deterministic tie rule). */
enum { CRT_BVH_HOST = 1, CRT_BVH_DEVICE = 2 };

enum { CRT_EXTEND_DEFAULT = 0, CRT_EXTEND_GRID_STRIDE = 1, CRT_EXTEND_PERSISTENT = 2, CRT_EXTEND_WAVE_CHUNK = 3 };

int crt_create(const crt_config* cfg, crt_context** out);
int crt_destroy(crt_context* ctx);
const char* crt_last_error(void);
int crt_device_count(int32_t* count);

/* Copies and uploads a scene (host BVH build + SoA upload). Limits per scene: 2^26 records of each
 * primitive type, and at most 4 GiB of BVH nodes and triangle, instance, sphere and quad records (about
 * 40M triangles with the default BVH); beyond them CRT_ERR_INVALID_SCENE. */
int crt_scene_upload(crt_context* ctx, const crt_scene_desc* desc, crt_scene_id* out);
int crt_scene_release(crt_context* ctx, crt_scene_id scene);
/* BVH statistics of an uploaded scene: nodes, leaves, depth, device bytes, build ms. */
typedef struct crt_scene_info {
    uint64_t num_nodes, num_leaves, num_prims, device_bytes;
    uint32_t max_depth, num_instances;
    uint32_t bvh_width, node_bytes;          /* children per node and bytes per node record */
    uint32_t leaf_size, builder;             /* max primitives per leaf the BVH was built with; CRT_BVH_* */
    double build_ms, upload_ms, sah_cost;
} crt_scene_info;
int crt_scene_get_info(crt_context* ctx, crt_scene_id scene, crt_scene_info* info);
/* Reads an uploaded scene's BVH and primitive records back from the device and checks the same
 * invariants as crt_scene_build_check (containment, every primitive record in exactly one leaf,
 * every node reachable) plus that the leaf-order records map one to one onto the scene's global
 * primitive ids. CRT_OK, or CRT_ERR_INTERNAL naming the violated invariant. Blocking. */
int crt_scene_verify(crt_context* ctx, crt_scene_id scene);
/* Host-only: flattens and builds a scene without a device, checks the BVH invariants
 * (containment, every primitive in exactly one leaf, reachability) and reports its statistics.
 * Same validation errors as crt_scene_upload (CRT_ERR_INVALID_SCENE for dangling indices). */
int crt_scene_build_check(const crt_scene_desc* desc, uint32_t leaf_size, uint32_t bvh_width, crt_scene_info* info);

/* Blocking render of one request (the launchGPU replacement). */
int crt_render(crt_context* ctx, crt_scene_id scene, const crt_render_desc* rd, crt_image* out,
               crt_stats* stats);

/* No counterpart in the reference: crt_render for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). It makes no
 * device allocation and calls no allocating library routine: its passes run in a fixed scratch
 * array that every GPU the library runs on holds for the life of the process (128 MiB; 256K live
 * paths, 1M path ids per pass, so a larger render takes more passes, never more memory). The scene
 * must be resident (crt_scene_upload). The caller owns the outputs, in device memory of the
 * context's device: accum_dev holds width*height*4 floats (rgb sums, sample count; zeroed here) and
 * rgb_dev (optional) width*height*3 bytes. The image equals crt_render's bit for bit. Keyed PCG
 * streams only (CRT_ERR_INVALID_ARGUMENT for CRT_RNG_MINSTD, CRT_RENDER_MEGAKERNEL or host
 * pointers). Blocking. Calls on one context are serialised; calls from different contexts on one
 * GPU share its scratch array and race. */
int crt_render_noalloc(crt_context* ctx, crt_scene_id scene, const crt_render_desc* rd, float* accum_dev,
                       uint8_t* rgb_dev, crt_stats* stats);

/* Closest-hit queries for arbitrary rays (first-hit parity). rays: n x 8 floats
 * (ox,oy,oz,tmin, dx,dy,dz,tmax). Outputs per ray: global primitive id (-1 = miss), instance
 * (-1 = none), t, and the barycentric / planar coordinates u, v. `flags` (CRT_TRACE_*):
 * CRT_TRACE_CAMERA hides invisible lights as the reference does for depth == max_depth;
 * CRT_TRACE_STAGE answers the queries with the render's closest-hit stage (the configured
 * extend_kernel, with its exact re-trace of undecided triangle tests) instead of the query kernel:
 * the interval is then the estimator's [1e-5, FLT_MAX) whatever tmin / tmax say, and a one-sided
 * light hit from behind is passed through as in the render (core/src/cameras/camera.cuh:179-183). */
enum { CRT_TRACE_CAMERA = 1, CRT_TRACE_STAGE = 2 };
int crt_trace_rays(crt_context* ctx, crt_scene_id scene, const float* rays, uint64_t n,
                   int32_t flags, int32_t* prim, int32_t* inst, float* t, float* uv);

/* Camera rays of sample 0 at jitter r = 0.5 with no lens/time draws (pixel-centre first hits). */
int crt_trace_primary(crt_context* ctx, crt_scene_id scene, uint32_t width, uint32_t height,
                      int32_t* prim, float* t);

/* ------------------------------------------------------------------ batched launch */
/* Renders several requests with fused wavefront passes over one scene table: every stage kernel
 * processes the rays of all requests together. Each request's image is identical to what
 * crt_render gives for it alone (keyed RNG + fixed per-pixel summation order). */
typedef struct crt_batch_item {
    crt_scene_id scene;
    crt_render_desc rd;
    crt_image out;
} crt_batch_item;
int crt_render_batch(crt_context* ctx, crt_batch_item* items, uint32_t n, crt_stats* stats);

/* ------------------------------------------------------------------ serving */
/* Continuous batching: requests are admitted against the HBM budget as they arrive, their
 * progressive passes are interleaved with those of requests already on the device, and each
 * completes independently. Thread-safe for concurrent submit/poll/wait/cancel on one context. */
int crt_submit(crt_context* ctx, crt_scene_id scene, const crt_render_desc* rd, crt_request_id* out);
int crt_poll(crt_context* ctx, crt_request_id req, int32_t* done);
int crt_wait(crt_context* ctx, crt_request_id req, crt_image* out, crt_stats* stats);
int crt_cancel(crt_context* ctx, crt_request_id req);

typedef struct crt_request_timing {
    double submit_s, admit_s, first_pass_s, done_s;  /* seconds on the library's steady clock */
    uint64_t rays;
    uint32_t passes;
    int32_t status;
} crt_request_timing;
int crt_request_get_timing(crt_context* ctx, crt_request_id req, crt_request_timing* out);

/* ------------------------------------------------------------------ autotuning */
typedef struct crt_tune_result {
    uint32_t max_paths;          /* live path slots */
    uint32_t grid_per_cu;        /* closest-hit workgroups per CU */
    uint32_t shade_grid_per_cu;  /* regenerate / shade workgroups per CU */
    uint32_t sub_passes;         /* concurrent sub-passes per pass */
    uint32_t extend_kernel;      /* CRT_EXTEND_* */
    uint32_t refill_min;         /* its parameter */
    uint32_t stack_lds;          /* LDS stack entries per lane */
    uint32_t bvh_width, leaf_size;   /* the scene's BVH after tuning */
    uint32_t trials;             /* trial renders timed */
    double best_mrays_s;         /* trial throughput of the choice */
    double initial_mrays_s;      /* trial throughput of the settings on entry */
} crt_tune_result;
/* No counterpart in the reference, which fixes its launch at 16x16-thread blocks
 * (core/src/gpu.cu:2072-2080). Picks the launch and memory shape of the streaming passes for
 * renders like `rd` of `scene` by timing trial renders (a sample range of `rd`, full image) of each
 * candidate, one tunable at a time: live path slots, closest-hit and shade workgroups per CU,
 * sub-passes, closest-hit work distribution, traversal stack depth in LDS. With `desc` (the
 * scene's description, as uploaded) it first rebuilds the scene's BVH at widths 2/4 and leaf sizes
 * 1/2/4 and keeps the fastest in place (same scene id). The choice becomes the context's setting
 * for later renders and uploads. Blocking; `desc` may be NULL. */
int crt_autotune(crt_context* ctx, crt_scene_id scene, const crt_scene_desc* desc, const crt_render_desc* rd,
                 crt_tune_result* out);

/* ------------------------------------------------------------------ mesh ingestion */
/* Wavefront OBJ reader (replaces the reference's tinyobjloader path, core/src/utilities/
 * mesh_loader.cuh:94-146 + core/src/gpu.cu:1776-1878): polygons are fanned into triangles, each
 * distinct (position, texcoord, normal) corner becomes one vertex, negative indices are resolved;
 * normals / uvs are NULL unless every corner has one. Host only; free with crt_obj_free. */
typedef struct crt_obj_mesh {
    float* positions; uint32_t num_vertices;
    uint32_t* indices; uint32_t num_triangles;
    float* normals;
    float* uvs;
} crt_obj_mesh;
int crt_obj_load(const char* path, crt_obj_mesh* out);
int crt_obj_free(crt_obj_mesh* mesh);

/* ------------------------------------------------------------------ scene files */
/* Reader of the reference's libconfig `.scene` files (core/src/scenes/scene_loader.cpp:21-120,
 * scene_builder, scene_manager::load_scene) for the render-path features: image, camera,
 * directional lights, solid/checker/PNG image textures, Lambertian/metal/glass materials, spheres,
 * quads, boxes and OBJ meshes. Everything else is counted in `unsupported` and skipped. The
 * descriptor's pointers stay valid until crt_scenefile_close. */
typedef struct crt_scenefile crt_scenefile;
typedef struct crt_scenefile_info {
    uint32_t width, height, spp, max_depth;   /* image {} values (the reference CLI overrides them) */
    float aspect_ratio;
    uint32_t unsupported;                     /* entries outside the render path, skipped */
    uint32_t skipped;                         /* primitives whose material name was not found */
    char output_path[512];
} crt_scenefile_info;
int crt_scenefile_load(const char* path, crt_scenefile** out);
int crt_scenefile_desc(crt_scenefile* file, crt_scene_desc* desc, crt_scenefile_info* info);
int crt_scenefile_close(crt_scenefile* file);

/* ------------------------------------------------------------------ named scene builder */
/* The reference's scene vocabulary (core/src/scenes/scene_builder.h, sceneConfig in
 * core/src/scenes/scene_config.h): textures and materials are added under names, primitives name
 * their material (and materials their texture), and the names are resolved as they are added, on
 * the host (the reference resolves them on the device, one thread comparing strings while it builds
 * the world: fetchTexture / fetchMaterial, core/src/gpu.cu:132-474). crt_scenefile_load builds
 * through these calls, and so does a sceneConfig-to-crt shim (one call per sceneConfig entry).
 * Names: NULL or "" adds an anonymous entry. A primitive whose material name is unknown is skipped
 * and counted (the reference's load_scene skips it, core/src/gpu.cu:531-534); a material whose
 * texture name is unknown is an error (CRT_ERR_INVALID_SCENE). The builder owns every array the
 * descriptor points to until crt_builder_destroy. Defaults: the scene builder's camera
 * (core/src/scenes/scene_builder.cpp:49: from (0, 0, 10) at the origin, y up, fov 70, focus 100,
 * aspect 16:9), a black background. */
typedef struct crt_builder crt_builder;

int crt_builder_create(crt_builder** out);
int crt_builder_destroy(crt_builder* b);
int crt_builder_camera(crt_builder* b, const crt_camera* camera);
int crt_builder_background(crt_builder* b, const float* rgb);
int crt_builder_solid_texture(crt_builder* b, const char* name, const float* rgb);
/* checker_texture (core/src/textures/checker_texture.cuh): each side a named texture, or a solid
 * colour when its name is NULL, empty or unknown (scene_factory.cuh:125-160) */
int crt_builder_checker_texture(crt_builder* b, const char* name, float scale, const char* even_texture,
                                const char* odd_texture, const float* even_rgb, const float* odd_rgb);
/* image_texture from a PNG or JPEG file; an unreadable file leaves an empty image (sampled cyan)
 * with a warning, as load_images prints and continues (core/src/gpu.cu:1728-1770) */
int crt_builder_image_texture(crt_builder* b, const char* name, const char* filepath);
int crt_builder_image_texture_pixels(crt_builder* b, const char* name, uint32_t width, uint32_t height,
                                     uint32_t channels, const uint8_t* pixels);
int crt_builder_lambertian(crt_builder* b, const char* name, const float* rgb, const char* texture);
int crt_builder_metal(crt_builder* b, const char* name, const float* rgb, float fuzz, const char* texture);
int crt_builder_dielectric(crt_builder* b, const char* name, float ior);
int crt_builder_isotropic(crt_builder* b, const char* name, const float* rgb, const char* texture);
/* oren_nayar as load_scene creates it (core/src/gpu.cu:440-451): only when `texture` names a known
 * texture, and then with the solid rgb (the reference's test is inverted); with an unknown or empty
 * texture name the reference's material has no texture and faults at its first hit: here it is not
 * created (its primitives are skipped) and counted unsupported */
int crt_builder_oren_nayar(crt_builder* b, const char* name, const float* rgb, const char* texture, float roughness,
                           float albedo_temp);
int crt_builder_diffuse_light(crt_builder* b, const char* name, const float* rgb, float intensity, int32_t one_sided);
/* 1 when a material of that name exists */
int crt_builder_has_material(crt_builder* b, const char* name);
/* lights (core/src/scenes/scene_builder.cpp:862-906 and the directional light): positions are
 * centres, colours linear RGB */
int crt_builder_directional_light(crt_builder* b, const char* name, const float* position, const float* u, const float* v,
                                  float intensity, const float* rgb, int32_t invisible);
int crt_builder_omni_light(crt_builder* b, const char* name, const float* position, float radius, float intensity,
                           const float* rgb, int32_t invisible);
int crt_builder_spot_light(crt_builder* b, const char* name, const float* position, const float* direction,
                           float cutoff_deg, float falloff, float intensity, float radius, const float* rgb,
                           int32_t invisible);
/* primitives; `mapping` and `xf` may be NULL */
int crt_builder_sphere(crt_builder* b, const char* name, const float* position, float radius, const char* material,
                       const crt_uvmap* mapping, const crt_transform* xf);
int crt_builder_quad(crt_builder* b, const char* name, const float* position, const float* u, const float* v,
                     const char* material, const crt_transform* xf);
/* box (core/src/primitives/box.cuh:41-74): [position - size/2, position + size/2] as 12 triangles */
int crt_builder_box(crt_builder* b, const char* name, const float* position, const float* size, const char* material,
                    const crt_transform* xf);
/* disk, cylinder or cone (CRT_SHAPE_*; createDisk / createCylinder / createCone,
 * core/src/scene_factory.cuh): `position` is the disk's centre or the base centre; the transform is
 * kept with the shape and applied at hit time as the reference's wrappers do */
int crt_builder_shape(crt_builder* b, uint32_t kind, const char* name, const float* position, float radius, float height,
                      const char* material, const crt_uvmap* mapping, const crt_transform* xf);
/* OBJ mesh (createObjMesh, core/src/scene_factory.cuh:373-385); use_smoothing = 0 drops the
 * file's normals (flat shading) */
int crt_builder_obj_mesh(crt_builder* b, const char* name, const char* filepath, const char* material,
                         int32_t use_smoothing);
/* counts n entries the caller left out as outside the render path */
int crt_builder_note_unsupported(crt_builder* b, uint32_t n);
/* The descriptor of everything added so far (valid until the next builder call), and the counts of
 * primitives skipped for an unknown material and of entries left out as unsupported. */
int crt_builder_desc(crt_builder* b, crt_scene_desc* out, uint32_t* skipped, uint32_t* unsupported);

/* ------------------------------------------------------------------ introspection */
/* Work counters of the last crt_render / crt_render_batch on this context. The traversal counts
 * (node_visits .. inst_visits) are filled only by a context created with count_work = 1. */
typedef struct crt_work_counters {
    uint64_t rays, retraces, light_pdf_tests;
    uint64_t node_visits, leaf_visits, tri_tests, sphere_tests, quad_tests, inst_visits;
    uint64_t extend_launches, passes;
    uint64_t iterations;        /* regenerate/extend/shade rounds of the streaming passes */
    uint64_t exact_queries;     /* closest-hit queries whose float triangle test could not decide a ray
                                   within rounding distance of an edge, re-traced in double */
} crt_work_counters;
int crt_last_counters(crt_context* ctx, crt_work_counters* out);

/* HBM accounting of a context: the admission budget, the bytes reserved by the context itself
 * (path-state pools, scene table, traversal overflow stacks), resident scenes and admitted requests. */
typedef struct crt_memory_info {
    uint64_t budget, reserved, scenes, requests;
} crt_memory_info;
int crt_memory_get_info(crt_context* ctx, crt_memory_info* out);

/* Test hook for the out-of-memory paths (no counterpart in the reference, whose allocation
 * failures exit(99), core/src/gpu.cu:119-129): the n-th device allocation the context makes from
 * now on (n = 0: the next one) fails with CRT_ERR_OUT_OF_MEMORY, once; n < 0 disarms it. Per-request
 * buffers reused from the context's cache are not allocations. */
int crt_debug_fail_alloc(crt_context* ctx, int32_t n);

/* VALU issue-rate calibration for the closest-hit kernel's roofline (no counterpart in the
 * reference): one launch of `waves_per_simd` waves on every SIMD of the context's GPU, each lane
 * running `iters` rounds of 16 independent chains of one instruction (kind 0: v_fma_f32, 1:
 * v_add_u32, 2: v_mul_f32; 3-7: v_fma_f32 with part of each wave active: lanes 0-15, every fourth
 * lane, lanes 0-31, lanes 0-62, lanes 0-47; 8: v_fma_f32 with no lane active; 9: v_cndmask_b32,
 * 10: v_max_f32, 11: v_cmp_lt_f32, 12: v_mov_b32, 13: v_cndmask_b32 with an SGPR-pair mask; 14 and
 * 15: v_fma_f32 with EXEC set to lanes 0-15 and 0-31, every lane alive), timed with HIP events after
 * one untimed launch. Returns the GPU time and the
 * wave instructions the rounds issue (waves x iters x 16). */
int crt_debug_valu_probe(crt_context* ctx, int32_t kind, uint32_t waves_per_simd, uint32_t iters, double* ms,
                         uint64_t* wave_insts);

/* Sizes of the public structs as compiled into the library (binding self-check). */
int crt_struct_sizes(uint64_t* sizes, uint32_t n);

#ifdef __cplusplus
}
#endif
#endif /* CORTEX_RT_H */
