// scenefile.cpp - reader of the reference's libconfig `.scene` files into a crt_scene_desc.
//
// The reference parses scenes with libconfig (core/src/scenes/scene_loader.cpp:21-120) into
// sceneConfig through scene_builder. This is a self-contained parser of the libconfig syntax
// (settings `name = value;` or `name : value`, groups {}, lists (), arrays [], strings with
// escapes and adjacent-string concatenation, ints/hex/floats/booleans, #, // and /* */ comments)
// and a mapping of the keys the render path uses:
//   image      width height maxDepth samplesPerPixel background{color} outputFilePath
//   camera     aspectRatio ("w:h" or number) lookFrom lookAt upAxis aperture focus fov
//   lights     directionals ( name position u v intensity color invisible active )
//   textures   solidColor (name color), checker (name scale oddColor/evenColor or
//              oddTextureName/evenTextureName), image (name filepath: PNG or JPEG)
//   materials  lambertian (name color | texture), metal (name color fuzziness), glass
//              (name refraction) and dielectric (same keys), isotropic (name color texture),
//              orennayar (name color texture albedo_temp roughness)
//   primitives spheres (name position radius material uvmapping), quads (name position u v
//              material), boxes (name position size material), cones / cylinders / disks (name
//              position radius height material uvmapping)
//   meshes     obj (name filepath material use_smoothing active)
// Semantics kept from the reference loader: colours written as integers are 0-255
// (getColor, scene_loader.cpp:1622-1641), metal reads `fuzziness` only (scene_loader.cpp:1128),
// box `position` is the centre (core/src/primitives/box.cuh:41-46), quad `position` is a corner,
// directional-light `position` is the centre, primitive transforms (translate, rotate, scale) are
// applied to spheres, quads, boxes and shapes and ignored for OBJ meshes (scene_factory.cuh:373-385). Fixed
// on purpose: counts are not truncated to 8 bits and `glass` materials are loaded (the reference
// sizes them from the `dielectric` key, scene_loader.cpp:191-192 vs 1092-1108, and drops them
// otherwise). Groups and the features outside the render path (omni/spot lights, other materials,
// textures and primitives, non-uniform scales of spheres) are counted in `unsupported` and skipped.
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
namespace cfg {

struct Value {
    enum Kind { NONE, BOOL, INT, FLOAT, STRING, GROUP, LIST, ARRAY } kind = NONE;
    bool b = false;
    long long i = 0;
    double f = 0.0;
    std::string s;
    std::vector<std::pair<std::string, std::unique_ptr<Value>>> members;  // GROUP
    std::vector<std::unique_ptr<Value>> items;                            // LIST / ARRAY
    const Value* get(const char* name) const {
        if (kind != GROUP) return nullptr;
        for (auto& m : members) if (m.first == name) return m.second.get();
        return nullptr;
    }
    bool number(double& out) const {
        if (kind == INT) { out = (double)i; return true; }
        if (kind == FLOAT) { out = f; return true; }
        return false;
    }
};

class Parser {
public:
    Parser(const std::string& text) : t_(text) {}
    bool parse(Value& root, std::string& err) {
        root.kind = Value::GROUP;
        if (!settings(root, 0)) { err = err_; return false; }
        skip();
        if (p_ < t_.size()) { err = where("unexpected character"); return false; }
        return true;
    }

private:
    const std::string& t_;
    size_t p_ = 0;
    std::string err_;

    std::string where(const std::string& what) {
        int line = 1;
        for (size_t k = 0; k < p_ && k < t_.size(); ++k) if (t_[k] == '\n') ++line;
        return "line " + std::to_string(line) + ": " + what;
    }
    bool fail(const std::string& what) { if (err_.empty()) err_ = where(what); return false; }
    void skip() {
        for (;;) {
            while (p_ < t_.size() && std::isspace((unsigned char)t_[p_])) ++p_;
            if (p_ < t_.size() && (t_[p_] == '#' || (t_[p_] == '/' && p_ + 1 < t_.size() && t_[p_ + 1] == '/'))) {
                while (p_ < t_.size() && t_[p_] != '\n') ++p_;
                continue;
            }
            if (p_ + 1 < t_.size() && t_[p_] == '/' && t_[p_ + 1] == '*') {
                size_t e = t_.find("*/", p_ + 2);
                p_ = e == std::string::npos ? t_.size() : e + 2;
                continue;
            }
            if (p_ < t_.size() && t_[p_] == '@') {  // @include and other directives: ignored
                while (p_ < t_.size() && t_[p_] != '\n') ++p_;
                continue;
            }
            return;
        }
    }
    bool name(std::string& out) {
        skip();
        size_t b = p_;
        if (p_ < t_.size() && (std::isalpha((unsigned char)t_[p_]) || t_[p_] == '*')) {
            ++p_;
            while (p_ < t_.size() && (std::isalnum((unsigned char)t_[p_]) || t_[p_] == '_' || t_[p_] == '-' || t_[p_] == '*')) ++p_;
        }
        if (p_ == b) return fail("setting name expected");
        out = t_.substr(b, p_ - b);
        return true;
    }
    bool settings(Value& g, int depth) {
        if (depth > 64) return fail("nesting too deep");
        for (;;) {
            skip();
            if (p_ >= t_.size() || t_[p_] == '}') return true;
            std::string n;
            if (!name(n)) return false;
            skip();
            if (p_ >= t_.size() || (t_[p_] != '=' && t_[p_] != ':')) return fail("'=' or ':' expected after " + n);
            ++p_;
            auto v = std::make_unique<Value>();
            if (!value(*v, depth + 1)) return false;
            skip();
            if (p_ < t_.size() && (t_[p_] == ';' || t_[p_] == ',')) ++p_;
            g.members.emplace_back(n, std::move(v));
        }
    }
    bool value(Value& v, int depth) {
        skip();
        if (p_ >= t_.size()) return fail("value expected");
        char c = t_[p_];
        if (c == '{') {
            ++p_;
            v.kind = Value::GROUP;
            if (!settings(v, depth)) return false;
            skip();
            if (p_ >= t_.size() || t_[p_] != '}') return fail("'}' expected");
            ++p_;
            return true;
        }
        if (c == '(' || c == '[') {
            char close = c == '(' ? ')' : ']';
            ++p_;
            v.kind = c == '(' ? Value::LIST : Value::ARRAY;
            for (;;) {
                skip();
                if (p_ < t_.size() && t_[p_] == close) { ++p_; return true; }
                auto it = std::make_unique<Value>();
                if (!value(*it, depth + 1)) return false;
                v.items.push_back(std::move(it));
                skip();
                if (p_ < t_.size() && t_[p_] == ',') { ++p_; continue; }
                if (p_ < t_.size() && t_[p_] == close) { ++p_; return true; }
                return fail(std::string("',' or '") + close + "' expected");
            }
        }
        if (c == '"') {
            v.kind = Value::STRING;
            while (p_ < t_.size() && t_[p_] == '"') {
                ++p_;
                while (p_ < t_.size() && t_[p_] != '"') {
                    char ch = t_[p_++];
                    if (ch == '\\' && p_ < t_.size()) {
                        char e = t_[p_++];
                        ch = e == 'n' ? '\n' : e == 't' ? '\t' : e == 'r' ? '\r' : e == 'f' ? '\f' : e;
                    }
                    v.s.push_back(ch);
                }
                if (p_ >= t_.size()) return fail("unterminated string");
                ++p_;
                skip();
            }
            return true;
        }
        size_t b = p_;
        while (p_ < t_.size() && (std::isalnum((unsigned char)t_[p_]) || t_[p_] == '.' || t_[p_] == '-' || t_[p_] == '+')) ++p_;
        std::string tok = t_.substr(b, p_ - b);
        if (tok.empty()) return fail("value expected");
        std::string low;
        for (char ch : tok) low.push_back((char)std::tolower((unsigned char)ch));
        if (low == "true" || low == "false") { v.kind = Value::BOOL; v.b = low == "true"; return true; }
        bool is_float = low.find_first_of(".e") != std::string::npos && low.rfind("0x", 0) != 0;
        char* end = nullptr;
        if (is_float) {
            v.kind = Value::FLOAT;
            v.f = std::strtod(tok.c_str(), &end);
        } else {
            v.kind = Value::INT;
            std::string digits = tok;
            while (!digits.empty() && (digits.back() == 'L' || digits.back() == 'l')) digits.pop_back();
            v.i = std::strtoll(digits.c_str(), &end, 0);
            if (end && *end == 0) return true;
            return fail("bad number " + tok);
        }
        if (!end || *end != 0) return fail("bad number " + tok);
        return true;
    }
};

}  // namespace cfg

// ------------------------------------------------------------------ scene mapping
// The file's entries go through the named builder (crt_builder_*), which resolves texture and
// material names and owns the descriptor's arrays.
struct SceneFile {
    std::string path, dir;
    crt_scenefile_info info{};
    crt_builder* b = nullptr;
    crt_scene_desc desc{};
    ~SceneFile() {
        if (b) crt_builder_destroy(b);
    }
};

static float num_of(const cfg::Value* v, float dflt) {
    double x;
    return (v && v->number(x)) ? (float)x : dflt;
}
static bool bool_of(const cfg::Value* v, bool dflt) {
    if (!v) return dflt;
    if (v->kind == cfg::Value::BOOL) return v->b;
    double x;
    if (v->number(x)) return x != 0.0;
    return dflt;
}
static std::string str_of(const cfg::Value* v) { return (v && v->kind == cfg::Value::STRING) ? v->s : std::string(); }
// getVector / getPoint: x, y, z as float or int (scene_loader.cpp:1560-1620)
static void vec_of(const cfg::Value* v, float* out, float dflt = 0.0f) {
    out[0] = num_of(v ? v->get("x") : nullptr, dflt);
    out[1] = num_of(v ? v->get("y") : nullptr, dflt);
    out[2] = num_of(v ? v->get("z") : nullptr, dflt);
}
// getColor: three integers are 0-255, anything else is read as floats (scene_loader.cpp:1622-1641)
static void color_of(const cfg::Value* v, float* out) {
    const cfg::Value* c[3] = {v ? v->get("r") : nullptr, v ? v->get("g") : nullptr, v ? v->get("b") : nullptr};
    bool ints = c[0] && c[1] && c[2] && c[0]->kind == cfg::Value::INT && c[1]->kind == cfg::Value::INT &&
                c[2]->kind == cfg::Value::INT && c[0]->i >= 0 && c[1]->i >= 0 && c[2]->i >= 0;
    for (int k = 0; k < 3; ++k) {
        if (ints) out[k] = (float)c[k]->i / 255.0f;
        else out[k] = (c[k] && c[k]->kind == cfg::Value::FLOAT) ? (float)c[k]->f : 0.0f;
    }
}
static const cfg::Value* list_of(const cfg::Value* g, const char* name) {
    const cfg::Value* v = g ? g->get(name) : nullptr;
    return (v && (v->kind == cfg::Value::LIST || v->kind == cfg::Value::ARRAY)) ? v : nullptr;
}

// Primitive transform: getTransform (scene_loader.cpp:1692-1704) reads translate, rotate (degrees
// about x, y, z) and scale (missing scale components 1); the builder applies it as the reference's
// wrappers see the primitive (crt_transform).
static const crt_transform* xform_of(const cfg::Value* it, crt_transform& x) {
    x = crt_transform{};
    const cfg::Value* tr = it ? it->get("transform") : nullptr;
    if (!tr || tr->kind != cfg::Value::GROUP) return nullptr;
    if (tr->get("translate")) { x.has |= CRT_XF_TRANSLATE; vec_of(tr->get("translate"), x.translate); }
    if (tr->get("rotate")) { x.has |= CRT_XF_ROTATE; vec_of(tr->get("rotate"), x.rotate_deg); }
    if (tr->get("scale")) { x.has |= CRT_XF_SCALE; vec_of(tr->get("scale"), x.scale, 1.0f); }
    return x.has ? &x : nullptr;
}

static std::string resolve(const SceneFile& sf, const std::string& p) {
    if (p.empty() || p[0] == '/') return p;
    std::string cand = sf.dir + "/" + p;
    if (FILE* f = std::fopen(cand.c_str(), "rb")) { std::fclose(f); return cand; }
    return p;
}

static int load_scenefile(SceneFile& sf, std::string& err) {
    FILE* f = std::fopen(sf.path.c_str(), "rb");
    if (!f) { err = "cannot open scene file " + sf.path; return CRT_ERR_IO; }
    std::string text;
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, n);
    std::fclose(f);
    size_t slash = sf.path.find_last_of('/');
    sf.dir = slash == std::string::npos ? "." : sf.path.substr(0, slash);
    cfg::Value root;
    cfg::Parser parser(text);
    if (!parser.parse(root, err)) { err = sf.path + ": " + err; return CRT_ERR_IO; }
    crt_scenefile_info& I = sf.info;
    // image (scene_loader.cpp:301-333); reference defaults from render_parameters.h:13-17
    I.width = 256; I.height = 144; I.spp = 100; I.max_depth = 100; I.aspect_ratio = 16.0f / 9.0f;
    float bg[3] = {0, 0, 0};
    if (const cfg::Value* im = root.get("image")) {
        I.width = (uint32_t)num_of(im->get("width"), (float)I.width);
        I.height = (uint32_t)num_of(im->get("height"), (float)I.height);
        I.max_depth = (uint32_t)num_of(im->get("maxDepth"), (float)I.max_depth);
        I.spp = (uint32_t)num_of(im->get("samplesPerPixel"), (float)I.spp);
        if (const cfg::Value* b = im->get("background")) {
            color_of(b->get("color"), bg);
            if (!str_of(b->get("filepath")).empty()) I.unsupported++;
        }
        std::string outp = str_of(im->get("outputFilePath"));
        std::snprintf(I.output_path, sizeof(I.output_path), "%s", outp.c_str());
    }
    // camera (scene_loader.cpp:335-366) over the scene builder's defaults (scene_builder.cpp:49: look
    // from (0, 0, 10) at the origin, y up, aperture 0, focus 100, perspective, fov 70, ortho height 0)
    crt_camera cam{};
    cam.lookfrom[2] = 10.0f;
    cam.vup[1] = 1.0f; cam.vfov_deg = 70.0f; cam.focus_dist = 100.0f; cam.time0 = 0.0f; cam.time1 = 1.0f;
    cam.aspect_ratio = I.aspect_ratio;
    if (const cfg::Value* c = root.get("camera")) {
        const cfg::Value* ar = c->get("aspectRatio");
        if (ar && ar->kind == cfg::Value::STRING) {
            double a = 0, b = 0;
            if (std::sscanf(ar->s.c_str(), "%lf:%lf", &a, &b) == 2 && a > 0 && b > 0) cam.aspect_ratio = (float)(a / b);
        } else if (ar) {
            cam.aspect_ratio = num_of(ar, cam.aspect_ratio);
        }
        if (c->get("lookFrom")) vec_of(c->get("lookFrom"), cam.lookfrom);
        if (c->get("lookAt")) vec_of(c->get("lookAt"), cam.lookat);
        if (c->get("upAxis")) vec_of(c->get("upAxis"), cam.vup);
        cam.aperture = num_of(c->get("aperture"), 0.0f);
        cam.focus_dist = num_of(c->get("focus"), cam.focus_dist);
        cam.vfov_deg = num_of(c->get("fov"), cam.vfov_deg);
        cam.orthographic = bool_of(c->get("orthographic"), false) ? 1u : 0u;
        cam.ortho_height = num_of(c->get("orthoHeight"), cam.ortho_height);
    }
    I.aspect_ratio = cam.aspect_ratio;
    int rc;
    if ((rc = crt_builder_create(&sf.b))) return rc;
    crt_builder* B = sf.b;
    if ((rc = crt_builder_camera(B, &cam)) || (rc = crt_builder_background(B, bg))) return rc;
    uint32_t unsupported = 0;
    // lights (scene_loader.cpp:276-300, 780-912): absent keys keep the loader's defaults (radius,
    // intensity, cutoff and falloff 0, colour black, invisible)
    if (const cfg::Value* L = root.get("lights")) {
        if (const cfg::Value* dl = list_of(L, "directionals")) {
            for (auto& it : dl->items) {
                if (!bool_of(it->get("active"), true)) continue;
                float pos[3], u[3], v[3], rgb[3] = {0, 0, 0};
                vec_of(it->get("position"), pos);
                vec_of(it->get("u"), u);
                vec_of(it->get("v"), v);
                color_of(it->get("color"), rgb);
                if ((rc = crt_builder_directional_light(B, str_of(it->get("name")).c_str(), pos, u, v,
                                                        num_of(it->get("intensity"), 0.0f), rgb,
                                                        bool_of(it->get("invisible"), true) ? 1 : 0)))
                    return rc;
            }
        }
        if (const cfg::Value* ol = list_of(L, "omnis")) {
            for (auto& it : ol->items) {
                if (!bool_of(it->get("active"), true)) continue;
                float pos[3] = {0, 0, 0}, rgb[3] = {0, 0, 0};
                if (it->get("position")) vec_of(it->get("position"), pos);
                if (it->get("color")) color_of(it->get("color"), rgb);
                if ((rc = crt_builder_omni_light(B, str_of(it->get("name")).c_str(), pos, num_of(it->get("radius"), 0.0f),
                                                 num_of(it->get("intensity"), 0.0f), rgb,
                                                 bool_of(it->get("invisible"), true) ? 1 : 0)))
                    return rc;
            }
        }
        if (const cfg::Value* sl = list_of(L, "spots")) {
            for (auto& it : sl->items) {
                if (!bool_of(it->get("active"), true)) continue;
                float pos[3] = {0, 0, 0}, dir[3] = {0, 0, 0}, rgb[3] = {0, 0, 0};
                if (it->get("position")) vec_of(it->get("position"), pos);
                if (it->get("direction")) vec_of(it->get("direction"), dir);
                if (it->get("color")) color_of(it->get("color"), rgb);
                if ((rc = crt_builder_spot_light(B, str_of(it->get("name")).c_str(), pos, dir, num_of(it->get("cutoff"), 0.0f),
                                                 num_of(it->get("falloff"), 0.0f), num_of(it->get("intensity"), 0.0f),
                                                 num_of(it->get("radius"), 0.0f), rgb,
                                                 bool_of(it->get("invisible"), true) ? 1 : 0)))
                    return rc;
            }
        }
    }
    // textures
    if (const cfg::Value* T = root.get("textures")) {
        if (const cfg::Value* l = list_of(T, "solidColor")) {
            for (auto& it : l->items) {
                float rgb[3];
                color_of(it->get("color"), rgb);
                if ((rc = crt_builder_solid_texture(B, str_of(it->get("name")).c_str(), rgb))) return rc;
            }
        }
        if (const cfg::Value* l = list_of(T, "image")) {
            for (auto& it : l->items) {
                const std::string fp = resolve(sf, str_of(it->get("filepath")));
                if ((rc = crt_builder_image_texture(B, str_of(it->get("name")).c_str(), fp.c_str()))) return rc;
            }
        }
        if (const cfg::Value* l = list_of(T, "checker")) {
            for (auto& it : l->items) {
                float even[3], odd[3];
                color_of(it->get("evenColor"), even);
                color_of(it->get("oddColor"), odd);
                if ((rc = crt_builder_checker_texture(B, str_of(it->get("name")).c_str(), num_of(it->get("scale"), 1.0f),
                                                      str_of(it->get("evenTextureName")).c_str(),
                                                      str_of(it->get("oddTextureName")).c_str(), even, odd)))
                    return rc;
            }
        }
        // textures outside the render path keep their names so materials still resolve; they are
        // sampled as mid grey and counted as unsupported
        const float grey[3] = {0.5f, 0.5f, 0.5f};
        for (const char* k : {"noise", "gradientColor", "bump", "normal", "displacement", "alpha", "emissive"}) {
            if (const cfg::Value* l = list_of(T, k)) {
                for (auto& it : l->items) {
                    if ((rc = crt_builder_solid_texture(B, str_of(it->get("name")).c_str(), grey))) return rc;
                    unsupported++;
                }
            }
        }
    }
    // materials
    if (const cfg::Value* M = root.get("materials")) {
        if (const cfg::Value* l = list_of(M, "lambertian")) {
            for (auto& it : l->items) {
                float rgb[3];
                color_of(it->get("color"), rgb);
                if ((rc = crt_builder_lambertian(B, str_of(it->get("name")).c_str(), rgb, str_of(it->get("texture")).c_str())))
                    return rc;
            }
        }
        if (const cfg::Value* l = list_of(M, "metal")) {
            for (auto& it : l->items) {
                float rgb[3];
                color_of(it->get("color"), rgb);
                if ((rc = crt_builder_metal(B, str_of(it->get("name")).c_str(), rgb, num_of(it->get("fuzziness"), 0.0f), nullptr)))
                    return rc;
            }
        }
        for (const char* k : {"glass", "dielectric"}) {
            if (const cfg::Value* l = list_of(M, k)) {
                for (auto& it : l->items) {
                    const std::string name = str_of(it->get("name"));
                    if (crt_builder_has_material(B, name.c_str()) == 1) continue;   // one material under both keys
                    if ((rc = crt_builder_dielectric(B, name.c_str(), num_of(it->get("refraction"), 1.0f)))) return rc;
                }
            }
        }
        // isotropic: added only when it names a texture, as addIsotropicMaterial does
        // (scene_loader.cpp: `if (!textureName.empty())`); a primitive naming one without is skipped
        if (const cfg::Value* l = list_of(M, "isotropic")) {
            for (auto& it : l->items) {
                const std::string tex = str_of(it->get("texture"));
                if (tex.empty()) continue;
                float rgb[3] = {0.0f, 0.0f, 0.0f};
                if (it->get("color")) color_of(it->get("color"), rgb);
                if ((rc = crt_builder_isotropic(B, str_of(it->get("name")).c_str(), rgb, tex.c_str()))) return rc;
            }
        }
        // oren_nayar: name color texture albedo_temp roughness (scene_loader.cpp:990-1016), created as
        // load_scene creates it (the builder call)
        if (const cfg::Value* l = list_of(M, "orennayar")) {
            for (auto& it : l->items) {
                float rgb[3] = {0.0f, 0.0f, 0.0f};
                if (it->get("color")) color_of(it->get("color"), rgb);
                if ((rc = crt_builder_oren_nayar(B, str_of(it->get("name")).c_str(), rgb, str_of(it->get("texture")).c_str(),
                                                 num_of(it->get("roughness"), 0.0f), num_of(it->get("albedo_temp"), 0.0f))))
                    return rc;
            }
        }
        // materials outside the render path become Lambertian with their colour (mid grey if none)
        for (const char* k : {"anisotropic", "phong"}) {
            if (const cfg::Value* l = list_of(M, k)) {
                for (auto& it : l->items) {
                    float rgb[3] = {0.5f, 0.5f, 0.5f};
                    if (it->get("color")) color_of(it->get("color"), rgb);
                    if ((rc = crt_builder_lambertian(B, str_of(it->get("name")).c_str(), rgb, nullptr))) return rc;
                    unsupported++;
                }
            }
        }
    }
    // primitives: a primitive whose material name is unknown is skipped by the builder
    // (fetchMaterial fails, gpu.cu:531-534)
    crt_transform xf;
    if (const cfg::Value* P = root.get("primitives")) {
        if (const cfg::Value* l = list_of(P, "spheres")) {
            for (auto& it : l->items) {
                if (!bool_of(it->get("active"), true)) continue;
                float c[3];
                vec_of(it->get("position"), c);
                crt_uvmap map{1, 1, 0, 0, 1, 1};
                if (const cfg::Value* uv = it->get("uvmapping")) {
                    map.scale_u = num_of(uv->get("scale_u"), 1.0f); map.scale_v = num_of(uv->get("scale_v"), 1.0f);
                    map.offset_u = num_of(uv->get("offset_u"), 0.0f); map.offset_v = num_of(uv->get("offset_v"), 0.0f);
                    map.repeat_u = num_of(uv->get("repeat_u"), 1.0f); map.repeat_v = num_of(uv->get("repeat_v"), 1.0f);
                }
                // a sphere maps to a sphere under translation, rotation and uniform scale; the texture
                // coordinates stay those of the world normal (the reference computes them in object space)
                if ((rc = crt_builder_sphere(B, str_of(it->get("name")).c_str(), c, num_of(it->get("radius"), 1.0f),
                                             str_of(it->get("material")).c_str(), &map, xform_of(it.get(), xf))))
                    return rc;
            }
        }
        if (const cfg::Value* l = list_of(P, "quads")) {
            for (auto& it : l->items) {
                if (!bool_of(it->get("active"), true)) continue;
                float c[3], u[3], v[3];
                vec_of(it->get("position"), c);
                vec_of(it->get("u"), u);
                vec_of(it->get("v"), v);
                if ((rc = crt_builder_quad(B, str_of(it->get("name")).c_str(), c, u, v, str_of(it->get("material")).c_str(),
                                           xform_of(it.get(), xf))))
                    return rc;
            }
        }
        if (const cfg::Value* l = list_of(P, "boxes")) {
            for (auto& it : l->items) {
                if (!bool_of(it->get("active"), true)) continue;
                float c[3], sz[3];
                vec_of(it->get("position"), c);
                vec_of(it->get("size"), sz);
                if ((rc = crt_builder_box(B, str_of(it->get("name")).c_str(), c, sz, str_of(it->get("material")).c_str(),
                                          xform_of(it.get(), xf))))
                    return rc;
            }
        }
        // cones, cylinders, disks in load_scene's order (core/src/gpu.cu:565-595), read as
        // addCone/addCylinderPrimitives/addDiskPrimitives read them (scene_loader.cpp:1311-1440):
        // position, radius and height default to 0, uvmapping to the unit mapping
        const struct { const char* key; uint32_t kind; } shapes[] = {
            {"cones", CRT_SHAPE_CONE}, {"cylinders", CRT_SHAPE_CYLINDER}, {"disks", CRT_SHAPE_DISK}};
        for (const auto& sk : shapes) {
            if (const cfg::Value* l = list_of(P, sk.key)) {
                for (auto& it : l->items) {
                    if (!bool_of(it->get("active"), true)) continue;
                    float c[3];
                    vec_of(it->get("position"), c);
                    crt_uvmap map{1, 1, 0, 0, 1, 1};
                    if (const cfg::Value* uv = it->get("uvmapping")) {
                        map.scale_u = num_of(uv->get("scale_u"), 1.0f); map.scale_v = num_of(uv->get("scale_v"), 1.0f);
                        map.offset_u = num_of(uv->get("offset_u"), 0.0f); map.offset_v = num_of(uv->get("offset_v"), 0.0f);
                        map.repeat_u = num_of(uv->get("repeat_u"), 1.0f); map.repeat_v = num_of(uv->get("repeat_v"), 1.0f);
                    }
                    if ((rc = crt_builder_shape(B, sk.kind, str_of(it->get("name")).c_str(), c, num_of(it->get("radius"), 0.0f),
                                                num_of(it->get("height"), 0.0f), str_of(it->get("material")).c_str(), &map,
                                                xform_of(it.get(), xf))))
                        return rc;
                }
            }
        }
        // tori are left out: torus::hit reports a hit without a distance (core/src/primitives/
        // torus.cuh:61-110, its intersection code commented out), which no renderer can reproduce
        for (const char* k : {"planes", "toruses", "volumes"})
            if (const cfg::Value* l = list_of(P, k)) unsupported += (uint32_t)l->items.size();
    }
    // meshes (OBJ)
    if (const cfg::Value* Ms = root.get("meshes")) {
        if (const cfg::Value* l = list_of(Ms, "obj")) {
            for (auto& it : l->items) {
                if (!bool_of(it->get("active"), true)) continue;
                const std::string fp = resolve(sf, str_of(it->get("filepath")));
                if ((rc = crt_builder_obj_mesh(B, str_of(it->get("name")).c_str(), fp.c_str(),
                                               str_of(it->get("material")).c_str(),
                                               bool_of(it->get("use_smoothing"), true) ? 1 : 0)))
                    return rc;
            }
        }
    }
    if ((rc = crt_builder_note_unsupported(B, unsupported))) return rc;
    uint32_t skipped = 0, unsup = 0;
    if ((rc = crt_builder_desc(B, &sf.desc, &skipped, &unsup))) return rc;
    I.skipped = skipped;
    I.unsupported += unsup;
    return CRT_OK;
}

}  // namespace crt

using namespace crt;

extern "C" {

int crt_scenefile_load(const char* path, crt_scenefile** out) {
    if (!path || !out) return set_error(CRT_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    auto sf = std::make_unique<SceneFile>();
    sf->path = path;
    std::string err;
    int rc = load_scenefile(*sf, err);
    if (rc) return err.empty() ? rc : set_error(rc, err);
    *out = reinterpret_cast<crt_scenefile*>(sf.release());
    return CRT_OK;
}

int crt_scenefile_desc(crt_scenefile* f, crt_scene_desc* desc, crt_scenefile_info* info) {
    if (!f) return set_error(CRT_ERR_INVALID_ARGUMENT, "null scene file");
    SceneFile* sf = reinterpret_cast<SceneFile*>(f);
    if (desc) *desc = sf->desc;
    if (info) *info = sf->info;
    return CRT_OK;
}

int crt_scenefile_close(crt_scenefile* f) {
    delete reinterpret_cast<SceneFile*>(f);
    return CRT_OK;
}

}  // extern "C"
