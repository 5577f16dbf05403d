// scene_builder.cpp - the named scene builder (crt_builder_*): the reference's scene_builder /
// sceneConfig vocabulary (materials and textures linked by name, primitives with transforms and
// uv mappings) resolved on the host into a crt_scene_desc.
//
// The reference keeps names in sceneConfig (core/src/scenes/scene_config.h:452-460) and resolves
// them on the device by string comparison while one GPU thread builds the world
// (strcmp_device + fetch*, core/src/gpu.cu:132-474). Here every name is resolved once, on the host,
// as it is added: the `.scene` reader (scenefile.cpp) and the launchGPU shim a maintainer links into
// the reference both go through these calls, so they resolve names the same way.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "cortex_rt.h"
#include "crt_runtime.h"
#include "scene_build.h"

namespace crt {

struct Builder {
    crt_scene_desc desc{};
    crt_camera camera{};
    float background[3] = {0, 0, 0};
    std::vector<crt_material> mats;
    std::vector<crt_texture> texs;
    std::vector<std::vector<uint8_t>> tex_pixels;
    std::vector<crt_area_light> lights;
    std::vector<crt_sphere_light> sphere_lights;
    std::vector<crt_sphere> spheres;
    std::vector<crt_quad> quads;
    std::vector<crt_shape> shapes;
    std::vector<crt_mesh> meshes;
    std::vector<std::unique_ptr<crt_obj_mesh, void (*)(crt_obj_mesh*)>> objs;
    std::vector<std::vector<float>> mesh_pos;
    std::vector<std::vector<uint32_t>> mesh_idx;
    std::map<std::string, int32_t> mat_by_name, tex_by_name;
    uint32_t skipped = 0, unsupported = 0;
    Builder() {
        // the reference scene builder's camera (core/src/scenes/scene_builder.cpp:49)
        camera.lookfrom[2] = 10.0f;
        camera.vup[1] = 1.0f; camera.vfov_deg = 70.0f; camera.focus_dist = 100.0f; camera.time1 = 1.0f;
        camera.aspect_ratio = 16.0f / 9.0f;
    }
};

namespace {

void free_obj(crt_obj_mesh* m) {
    if (m) { crt_obj_free(m); delete m; }
}

void copy3(float* d, const float* s, float dflt = 0.0f) {
    for (int k = 0; k < 3; ++k) d[k] = s ? s[k] : dflt;
}

int32_t add_texture(Builder& b, const char* name, const crt_texture& t) {
    const int32_t idx = (int32_t)b.texs.size();
    if (name && *name) b.tex_by_name[name] = idx;
    b.texs.push_back(t);
    b.tex_pixels.emplace_back();
    return idx;
}

int32_t add_material(Builder& b, const char* name, const crt_material& m) {
    const int32_t idx = (int32_t)b.mats.size();
    if (name && *name) b.mat_by_name[name] = idx;
    b.mats.push_back(m);
    return idx;
}

// fetchMaterial (core/src/gpu.cu:407-474): an unknown name skips the primitive (gpu.cu:531-534)
bool material_of(Builder& b, const char* name, uint32_t& out) {
    auto it = b.mat_by_name.find(name ? name : "");
    if (it == b.mat_by_name.end()) { b.skipped++; return false; }
    out = (uint32_t)it->second;
    return true;
}

int texture_of(Builder& b, const char* name, int32_t& out) {
    out = -1;
    if (!name || !*name) return CRT_OK;
    auto it = b.tex_by_name.find(name);
    if (it == b.tex_by_name.end()) return set_error(CRT_ERR_INVALID_SCENE, std::string("texture '") + name + "' not found");
    out = it->second;
    return CRT_OK;
}

}  // namespace

// Primitive transform (rt::transform, core/src/misc/transform.cuh; applyTransform,
// core/src/scene_factory.cuh:710-725): translate, then rotate (degrees about x, y, z), then scale.
// A world ray reaches object space divided by the scale (scale.cuh:50-56), rotated by
// Rx(-x) Ry(-y) Rz(-z) (rotate.cuh:104-119; glm::rotate post-multiplies) and offset by -T
// (translate.cuh:38-43), so the surface it hits is the primitive mapped by
// p -> S Rz(z) Ry(y) Rx(x) (p + T). That map is applied to the geometry. (With one rotation axis it
// is also the map of the reported hit point; with two or more the reference reports its hit point
// through Rx Ry Rz instead, off the ray.)
struct XformMap {
    bool any = false, uniform = true;
    float T[3] = {0, 0, 0}, L[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    float scale = 1.0f;
    explicit XformMap(const crt_transform* x) {
        if (!x || !(x->has & 7u)) return;
        any = true;
        float t[3] = {0, 0, 0}, r[3] = {0, 0, 0}, s[3] = {1, 1, 1};
        if (x->has & CRT_XF_TRANSLATE) copy3(t, x->translate);
        if (x->has & CRT_XF_ROTATE) copy3(r, x->rotate_deg);
        if (x->has & CRT_XF_SCALE) copy3(s, x->scale);
        std::memcpy(T, t, sizeof(t));
        const float kPi = 3.14159265358979323846f;   // degrees_to_radians in float (math_utils.cuh:55-58)
        float M[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        for (int axis = 2; axis >= 0; --axis) {       // M = Rz(z) Ry(y) Rx(x)
            const float a = r[axis] * kPi / 180.0f, c = std::cos(a), sn = std::sin(a);
            float R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
            const int i = (axis + 1) % 3, j = (axis + 2) % 3;
            R[3 * i + i] = c; R[3 * i + j] = -sn; R[3 * j + i] = sn; R[3 * j + j] = c;
            float P[9];
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q) P[3 * p + q] = M[3 * p] * R[q] + M[3 * p + 1] * R[3 + q] + M[3 * p + 2] * R[6 + q];
            std::memcpy(M, P, sizeof(P));
        }
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) L[3 * a + c] = s[a] * M[3 * a + c];
        uniform = std::fabs(s[0]) == std::fabs(s[1]) && std::fabs(s[1]) == std::fabs(s[2]);
        scale = std::fabs(s[0]);
    }
    void linear(const float* v, float* out) const {
        float r[3];
        for (int a = 0; a < 3; ++a) r[a] = L[3 * a] * v[0] + L[3 * a + 1] * v[1] + L[3 * a + 2] * v[2];
        std::memcpy(out, r, sizeof(r));
    }
    void point(const float* p, float* out) const {
        const float q[3] = {p[0] + T[0], p[1] + T[1], p[2] + T[2]};
        linear(q, out);
    }
};

}  // namespace crt

using namespace crt;

#define CRT_B(b) Builder& B = *reinterpret_cast<Builder*>(b)
#define CRT_B_CHECK(b) if (!(b)) return set_error(CRT_ERR_INVALID_ARGUMENT, "null builder")

extern "C" {

int crt_builder_create(crt_builder** out) {
    if (!out) return set_error(CRT_ERR_INVALID_ARGUMENT, "null argument");
    *out = reinterpret_cast<crt_builder*>(new Builder());
    return CRT_OK;
}

int crt_builder_destroy(crt_builder* b) {
    delete reinterpret_cast<Builder*>(b);
    return CRT_OK;
}

int crt_builder_camera(crt_builder* b, const crt_camera* cam) {
    CRT_B_CHECK(b);
    if (!cam) return set_error(CRT_ERR_INVALID_ARGUMENT, "null camera");
    CRT_B(b);
    B.camera = *cam;
    return CRT_OK;
}

int crt_builder_background(crt_builder* b, const float* rgb) {
    CRT_B_CHECK(b);
    CRT_B(b);
    copy3(B.background, rgb);
    return CRT_OK;
}

int crt_builder_solid_texture(crt_builder* b, const char* name, const float* rgb) {
    CRT_B_CHECK(b);
    CRT_B(b);
    crt_texture t{};
    t.type = CRT_TEX_SOLID; t.even = t.odd = -1;
    copy3(t.color, rgb);
    add_texture(B, name, t);
    return CRT_OK;
}

// checker_texture(scale, even, odd) (core/src/textures/checker_texture.cuh:23-31): each side a
// named texture or, when the name is empty or unknown, a solid colour (scene_factory.cuh:125-160)
int crt_builder_checker_texture(crt_builder* b, const char* name, float scale, const char* even_texture,
                                const char* odd_texture, const float* even_rgb, const float* odd_rgb) {
    CRT_B_CHECK(b);
    CRT_B(b);
    auto side = [&](const char* tn, const float* rgb) -> int32_t {
        if (tn && *tn) {
            auto it = B.tex_by_name.find(tn);
            if (it != B.tex_by_name.end()) return it->second;
        }
        crt_texture s{};
        s.type = CRT_TEX_SOLID; s.even = s.odd = -1;
        copy3(s.color, rgb);
        return add_texture(B, nullptr, s);
    };
    crt_texture t{};
    t.type = CRT_TEX_CHECKER;
    t.scale = scale;
    t.odd = side(odd_texture, odd_rgb);
    t.even = side(even_texture, even_rgb);
    add_texture(B, name, t);
    return CRT_OK;
}

// image_texture from a file (load_images, core/src/gpu.cu:1728-1770): PNG or JPEG. A file that
// cannot be read leaves an empty image, sampled as cyan, with a warning, as the reference prints
// and continues.
int crt_builder_image_texture(crt_builder* b, const char* name, const char* filepath) {
    CRT_B_CHECK(b);
    CRT_B(b);
    crt_texture t{};
    t.type = CRT_TEX_IMAGE; t.even = t.odd = -1; t.channels = 3;
    std::vector<uint8_t> px;
    std::string err;
    if (!filepath || !read_image(filepath, px, t.width, t.height, t.channels, err)) {
        std::fprintf(stderr, "[WARNING] %s\n", filepath ? err.c_str() : "image texture without a file path");
        t.width = t.height = 0; t.channels = 3;
        px.clear();
    }
    add_texture(B, name, t);
    B.tex_pixels.back() = std::move(px);
    return CRT_OK;
}

int crt_builder_image_texture_pixels(crt_builder* b, const char* name, uint32_t width, uint32_t height, uint32_t channels,
                                     const uint8_t* pixels) {
    CRT_B_CHECK(b);
    CRT_B(b);
    if ((channels != 3 && channels != 4) || (width && height && !pixels))
        return set_error(CRT_ERR_INVALID_ARGUMENT, "image texture needs 3 or 4 channels and pixels");
    crt_texture t{};
    t.type = CRT_TEX_IMAGE; t.even = t.odd = -1;
    t.width = width; t.height = height; t.channels = channels;
    add_texture(B, name, t);
    if (pixels) B.tex_pixels.back().assign(pixels, pixels + (size_t)width * height * channels);
    return CRT_OK;
}

int crt_builder_lambertian(crt_builder* b, const char* name, const float* rgb, const char* texture) {
    CRT_B_CHECK(b);
    CRT_B(b);
    crt_material m{};
    m.type = CRT_MAT_LAMBERTIAN;
    copy3(m.albedo, rgb);
    int rc = texture_of(B, texture, m.texture);
    if (rc) return rc;
    add_material(B, name, m);
    return CRT_OK;
}

// isotropic (core/src/materials/isotropic.cuh): the albedo of its texture, or of a solid rgb when no
// texture is named (createIsotropicMaterial, core/src/gpu.cu:407-421)
int crt_builder_isotropic(crt_builder* b, const char* name, const float* rgb, const char* texture) {
    CRT_B_CHECK(b);
    CRT_B(b);
    crt_material m{};
    m.type = CRT_MAT_ISOTROPIC;
    copy3(m.albedo, rgb);
    int rc = texture_of(B, texture, m.texture);
    if (rc) return rc;
    add_material(B, name, m);
    return CRT_OK;
}

int crt_builder_oren_nayar(crt_builder* b, const char* name, const float* rgb, const char* texture, float roughness,
                           float albedo_temp) {
    CRT_B_CHECK(b);
    CRT_B(b);
    if (!texture || !texture[0] || !B.tex_by_name.count(texture)) {
        B.unsupported++;
        return CRT_OK;
    }
    crt_material m{};
    m.type = CRT_MAT_OREN_NAYAR;
    copy3(m.albedo, rgb);
    m.texture = -1;
    m.fuzz = roughness;
    m.ior = albedo_temp;
    add_material(B, name, m);
    return CRT_OK;
}

int crt_builder_metal(crt_builder* b, const char* name, const float* rgb, float fuzz, const char* texture) {
    CRT_B_CHECK(b);
    CRT_B(b);
    crt_material m{};
    m.type = CRT_MAT_METAL;
    copy3(m.albedo, rgb);
    m.fuzz = fuzz;
    int rc = texture_of(B, texture, m.texture);
    if (rc) return rc;
    add_material(B, name, m);
    return CRT_OK;
}

int crt_builder_dielectric(crt_builder* b, const char* name, float ior) {
    CRT_B_CHECK(b);
    CRT_B(b);
    crt_material m{};
    m.type = CRT_MAT_DIELECTRIC; m.texture = -1;
    m.albedo[0] = m.albedo[1] = m.albedo[2] = 1.0f;
    m.ior = ior;
    add_material(B, name, m);
    return CRT_OK;
}

int crt_builder_diffuse_light(crt_builder* b, const char* name, const float* rgb, float intensity, int32_t one_sided) {
    CRT_B_CHECK(b);
    CRT_B(b);
    crt_material m{};
    m.type = CRT_MAT_DIFFUSE_LIGHT; m.texture = -1;
    copy3(m.albedo, rgb);
    m.intensity = intensity;
    m.flags = one_sided ? CRT_MATF_ONE_SIDED : 0u;
    add_material(B, name, m);
    return CRT_OK;
}

// directional_light (core/src/lights/directional_light.cuh:45-172): a quad centred at `position`
int crt_builder_directional_light(crt_builder* b, const char* name, const float* position, const float* u, const float* v,
                                  float intensity, const float* rgb, int32_t invisible) {
    CRT_B_CHECK(b);
    CRT_B(b);
    (void)name;
    crt_area_light a{};
    copy3(a.center, position); copy3(a.u, u); copy3(a.v, v); copy3(a.rgb, rgb);
    a.intensity = intensity;
    a.invisible = invisible ? 1u : 0u;
    B.lights.push_back(a);
    return CRT_OK;
}

int crt_builder_has_material(crt_builder* b, const char* name) {
    CRT_B_CHECK(b);
    CRT_B(b);
    return B.mat_by_name.count(name ? name : "") ? 1 : 0;
}

// omni_light (core/src/lights/omni_light.cuh; addOmniDirectionalLight, scene_builder.cpp:862-883)
int crt_builder_omni_light(crt_builder* b, const char* name, const float* position, float radius, float intensity,
                           const float* rgb, int32_t invisible) {
    CRT_B_CHECK(b);
    CRT_B(b);
    (void)name;
    crt_sphere_light s{};
    s.kind = CRT_LIGHT_OMNI;
    copy3(s.center, position); copy3(s.rgb, rgb);
    s.radius = radius;
    s.intensity = intensity;
    s.invisible = invisible ? 1u : 0u;
    B.sphere_lights.push_back(s);
    return CRT_OK;
}

// spot_light (core/src/lights/spot_light.cuh; addSpotLight, scene_builder.cpp:885-906)
int crt_builder_spot_light(crt_builder* b, const char* name, const float* position, const float* direction,
                           float cutoff_deg, float falloff, float intensity, float radius, const float* rgb,
                           int32_t invisible) {
    CRT_B_CHECK(b);
    CRT_B(b);
    (void)name;
    crt_sphere_light s{};
    s.kind = CRT_LIGHT_SPOT;
    copy3(s.center, position); copy3(s.direction, direction); copy3(s.rgb, rgb);
    s.cutoff_deg = cutoff_deg;
    s.falloff = falloff;
    s.intensity = intensity;
    s.radius = radius;
    s.invisible = invisible ? 1u : 0u;
    B.sphere_lights.push_back(s);
    return CRT_OK;
}

// sphere (core/src/primitives/sphere.cuh) with its uv mapping; under a transform the centre moves
// and a uniform scale scales the radius (a non-uniform scale, an ellipsoid, is counted unsupported
// and left out of the radius)
int crt_builder_sphere(crt_builder* b, const char* name, const float* position, float radius, const char* material,
                       const crt_uvmap* mapping, const crt_transform* xf) {
    CRT_B_CHECK(b);
    CRT_B(b);
    (void)name;
    crt_sphere s{};
    if (!material_of(B, material, s.material)) return CRT_OK;
    copy3(s.center, position);
    s.radius = radius;
    s.mapping = mapping ? *mapping : crt_uvmap{1, 1, 0, 0, 1, 1};
    const XformMap x(xf);
    if (x.any) {
        float c[3];
        x.point(s.center, c);
        std::memcpy(s.center, c, sizeof(c));
        if (x.uniform) s.radius *= x.scale; else B.unsupported++;
    }
    B.spheres.push_back(s);
    return CRT_OK;
}

// quad (core/src/primitives/quad.cuh): `position` is a corner, u and v its edges
int crt_builder_quad(crt_builder* b, const char* name, const float* position, const float* u, const float* v,
                     const char* material, const crt_transform* xf) {
    CRT_B_CHECK(b);
    CRT_B(b);
    (void)name;
    crt_quad q{};
    if (!material_of(B, material, q.material)) return CRT_OK;
    copy3(q.corner, position); copy3(q.u, u); copy3(q.v, v);
    const XformMap x(xf);
    if (x.any) {
        float c[3], uu[3], vv[3];
        x.point(q.corner, c); x.linear(q.u, uu); x.linear(q.v, vv);
        std::memcpy(q.corner, c, sizeof(c)); std::memcpy(q.u, uu, sizeof(uu)); std::memcpy(q.v, vv, sizeof(vv));
    }
    B.quads.push_back(q);
    return CRT_OK;
}

// box (core/src/primitives/box.cuh:41-74): the six faces of [position - size/2, position + size/2],
// here as 12 triangles
int crt_builder_box(crt_builder* b, const char* name, const float* position, const float* size, const char* material,
                    const crt_transform* xf) {
    CRT_B_CHECK(b);
    CRT_B(b);
    (void)name;
    uint32_t m;
    if (!material_of(B, material, m)) return CRT_OK;
    float c[3], s[3];
    copy3(c, position); copy3(s, size);
    std::vector<float> pos(24);
    for (int k = 0; k < 8; ++k)
        for (int a = 0; a < 3; ++a) pos[3 * k + a] = (k >> a & 1) ? c[a] + s[a] / 2.0f : c[a] - s[a] / 2.0f;
    const XformMap x(xf);
    if (x.any)
        for (int k = 0; k < 8; ++k) {
            float w[3];
            x.point(&pos[3 * k], w);
            std::memcpy(&pos[3 * k], w, sizeof(w));
        }
    B.mesh_pos.push_back(std::move(pos));
    B.mesh_idx.push_back({0, 1, 3, 0, 3, 2, 4, 6, 7, 4, 7, 5, 0, 4, 5, 0, 5, 1, 2, 3, 7, 2, 7, 6,
                          0, 2, 6, 0, 6, 4, 1, 5, 7, 1, 7, 3});
    crt_mesh me{};
    me.material = m;
    me.num_vertices = 8;
    me.num_triangles = 12;
    B.meshes.push_back(me);
    return CRT_OK;
}

// disk / cylinder / cone (core/src/primitives/disk.cuh, cylinder.cuh, cone.cuh): the transform stays
// with the shape, whose hit runs in its own frame as under the reference's wrappers
int crt_builder_shape(crt_builder* b, uint32_t kind, const char* name, const float* position, float radius, float height,
                      const char* material, const crt_uvmap* mapping, const crt_transform* xf) {
    CRT_B_CHECK(b);
    CRT_B(b);
    (void)name;
    if (kind > CRT_SHAPE_CONE) return set_error(CRT_ERR_INVALID_ARGUMENT, "unknown shape kind");
    crt_shape s{};
    if (!material_of(B, material, s.material)) return CRT_OK;
    s.kind = kind;
    copy3(s.center, position);
    s.radius = radius;
    s.height = height;
    s.mapping = mapping ? *mapping : crt_uvmap{1, 1, 0, 0, 1, 1};
    if (xf) s.transform = *xf;
    B.shapes.push_back(s);
    return CRT_OK;
}

// OBJ mesh (createObjMesh, core/src/scene_factory.cuh:373-385): position and transform are not
// applied, as in the reference; use_smoothing = 0 keeps flat normals
int crt_builder_obj_mesh(crt_builder* b, const char* name, const char* filepath, const char* material,
                         int32_t use_smoothing) {
    CRT_B_CHECK(b);
    CRT_B(b);
    (void)name;
    uint32_t m;
    if (!material_of(B, material, m)) return CRT_OK;
    std::unique_ptr<crt_obj_mesh, void (*)(crt_obj_mesh*)> om(new crt_obj_mesh(), free_obj);
    std::memset(om.get(), 0, sizeof(crt_obj_mesh));
    if (!filepath || crt_obj_load(filepath, om.get()) != CRT_OK)
        return set_error(CRT_ERR_IO, std::string("cannot load mesh ") + (filepath ? filepath : "(null)"));
    if (!use_smoothing) { std::free(om->normals); om->normals = nullptr; }
    crt_mesh me{};
    me.positions = om->positions; me.num_vertices = om->num_vertices;
    me.indices = om->indices; me.num_triangles = om->num_triangles;
    me.normals = om->normals; me.uvs = om->uvs; me.material = m;
    B.mesh_pos.emplace_back();
    B.mesh_idx.emplace_back();
    B.meshes.push_back(me);
    B.objs.push_back(std::move(om));
    return CRT_OK;
}

int crt_builder_desc(crt_builder* b, crt_scene_desc* out, uint32_t* skipped, uint32_t* unsupported) {
    CRT_B_CHECK(b);
    CRT_B(b);
    if (!out) return set_error(CRT_ERR_INVALID_ARGUMENT, "null argument");
    // boxes own their vertex arrays (vectors no longer move now)
    for (size_t k = 0; k < B.meshes.size(); ++k)
        if (!B.mesh_pos[k].empty()) { B.meshes[k].positions = B.mesh_pos[k].data(); B.meshes[k].indices = B.mesh_idx[k].data(); }
    for (size_t k = 0; k < B.texs.size(); ++k)
        if (B.texs[k].type == CRT_TEX_IMAGE) B.texs[k].pixels = B.tex_pixels[k].empty() ? nullptr : B.tex_pixels[k].data();
    crt_scene_desc& d = B.desc;
    std::memset(&d, 0, sizeof(d));
    d.abi_version = CRT_ABI_VERSION;
    d.camera = B.camera;
    copy3(d.background, B.background);
    d.materials = B.mats.data(); d.num_materials = (uint32_t)B.mats.size();
    d.textures = B.texs.data(); d.num_textures = (uint32_t)B.texs.size();
    d.lights = B.lights.data(); d.num_lights = (uint32_t)B.lights.size();
    d.sphere_lights = B.sphere_lights.data(); d.num_sphere_lights = (uint32_t)B.sphere_lights.size();
    d.spheres = B.spheres.data(); d.num_spheres = (uint32_t)B.spheres.size();
    d.quads = B.quads.data(); d.num_quads = (uint32_t)B.quads.size();
    d.shapes = B.shapes.data(); d.num_shapes = (uint32_t)B.shapes.size();
    d.meshes = B.meshes.data(); d.num_meshes = (uint32_t)B.meshes.size();
    *out = d;
    if (skipped) *skipped = B.skipped;
    if (unsupported) *unsupported = B.unsupported;
    return CRT_OK;
}

int crt_builder_note_unsupported(crt_builder* b, uint32_t n) {
    CRT_B_CHECK(b);
    CRT_B(b);
    B.unsupported += n;
    return CRT_OK;
}

}  // extern "C"
